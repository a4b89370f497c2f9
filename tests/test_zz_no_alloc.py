"""fluid_gpu_diagnostics_noalloc: the diagnostics of src/logging.c:40-172
with their partial sums in the library's static device scratch instead of
context memory.  Through the C-ABI, its results equal the synchronising
entry points (fluid_gpu_divergence / momentum / energy) bit for bit, on two
calls in a row of which the second has a smaller input (the scratch keeps
the first call's contents and must not leak into the second)."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")
torch = pytest.importorskip("torch")


def _check(nx, ny, ly, seed):
    g = eif.Solver(itot=nx, jtot=ny, ly=ly, Re=100.0)
    ux, uy, p = cases.random_fields(nx, ny, seed, amp=2.0)
    g.upload(ux, uy, p)
    g.update_boundaries()
    ref = [g.divergence(), tuple(g.momentum()), tuple(g.energy())]
    out = torch.full((3, 2), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for which in range(3):
        g.diagnostics_noalloc(which, out[which].data_ptr())
    g.synchronize()
    got = out.cpu().numpy()
    for which in range(3):
        assert got[which].tolist() == list(ref[which]), (nx, ny, which, got[which], ref[which])
    assert ref[0][0] > 0.0 and abs(ref[1][0]) > 0.0
    g.close()


def test_noalloc_diagnostics_match_the_synchronising_calls():
    print("noalloc diagnostics: 256x384 then 40x24", flush=True)
    _check(256, 384, 1.5, 3)
    _check(40, 24, 1.0, 4)


def test_noalloc_rejects_bad_arguments():
    g = eif.Solver(itot=16, jtot=16)
    with pytest.raises(eif.EIFError, match="-1"):
        g.diagnostics_noalloc(3, 8)
    with pytest.raises(eif.EIFError, match="-1"):
        g.diagnostics_noalloc(0, 0)
    g.close()
