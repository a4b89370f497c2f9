"""bench.py's command line without a GPU: more than one GPU is refused with
a message, and the script's constants describe the measured workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line for a refused run


def test_workload_constants():
    src = open(os.path.join(ROOT, "bench.py")).read()
    ns = {}
    # the module-level constants only (no GPU work at import)
    for line in src.splitlines():
        if line.startswith(("NX, NY", "NP =", "PROFILE_EVERY =", "PEAK_FP64 =")):
            exec(line, ns)
    assert (ns["NX"], ns["NY"], ns["NP"]) == (2048, 4096, 1024)
    assert ns["PEAK_FP64"] == 78.6e12
    assert ns["PROFILE_EVERY"] > 1
