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
