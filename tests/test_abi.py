"""The C-ABI library loads and exports every function include/eif_gpu.h
declares; without a device, creation fails with EIF_ERR_NODEVICE instead
of falling back to anything else."""
import ctypes
import os
import re

import pytest

import ellipsesinflows_amd as eif

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eif_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("fluid_gpu_create", "poisson_gpu_compute_potential",
                     "suspensions_gpu_compute_inertia", "eif_gpu_step"):
        assert required in names
    assert set(names) == set(eif.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = eif.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_product_library_does_not_link_the_oracle():
    with open(eif.library_path, "rb") as f:
        data = f.read()
    assert b"liboracle" not in data and b"orc_" not in data


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a device is present")
def test_create_without_device_reports_nodevice():
    with pytest.raises(eif.EIFError, match="-5"):
        eif.Solver(itot=16, jtot=16)
    lib = eif.load_library()
    assert b"HIP device" in lib.eif_gpu_last_error()


def test_partition_layout_fits_the_reduced_system_in_lds():
    """The partitioned y-solve's reduced system keeps 26 P doubles per
    workgroup in LDS (k_part_reduced): the chunk count stays within 64 KiB,
    and lines it cannot split that way report (0, 0) -- the context then
    uses the Thomas sweep -- instead of failing at launch."""
    assert eif.partition_layout(4096) == (64, 64)
    assert eif.partition_layout(1024) == (64, 16)
    assert eif.partition_layout(12) == (0, 0)
    for jtot in (8 * 789, 8 * 1024, 64 * 315, 64 * 316, 6000, 4097, 3, 16):
        L, P = eif.partition_layout(jtot)
        if L:
            assert L * P == jtot and 2 <= P and 26 * P * 8 <= 65536
    assert eif.partition_layout(8 * 789) == (0, 0)
