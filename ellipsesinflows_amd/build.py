"""Build recipes for the native libraries (run on a machine with hipcc).

``build_gpu()`` compiles the HIP sources into ``ellipsesinflows_amd/libeif_gpu.so``
for gfx950; ``build_host()`` compiles the host modules (parameters, npy
save/restart, logging) into ``libeif_host.so`` and the C driver
``eif_gpu_main`` (a drop-in for the reference's ``a.out``) against
``libeif_gpu.so``.  Everything builds in-tree.
"""
import concurrent.futures
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libeif_gpu.so")
EXE = os.path.join(HERE, "eif_gpu_main")
HOSTLIB = os.path.join(HERE, "libeif_host.so")

HIP_SOURCES = ["eif_api.hip", "eif_fluid.hip", "eif_poisson.hip", "eif_suspensions.hip",
               "eif_statistics.hip", "eif_adjoint.hip"]
HOST_SOURCES = ["param.c", "npyio.c", "save.c", "logging.c"]
HOST_CFLAGS = ["-O2", "-std=c99", "-Wall", "-Wextra", "-Werror"]

# -ffp-contract=off: expressions round like the reference's ISO C99 build.
# packed FP32 instructions are disabled on this platform.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), proc.stdout))
    return proc.stdout


def build_gpu(hipcc="hipcc", jobs=None):
    """Each HIP source to an object in parallel (in a temporary directory),
    then one shared library."""
    jobs = jobs or min(len(HIP_SOURCES), os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(prefix="eif_build_") as tmp:
        def compile_one(src):
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            _run([hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", src,
                                          "-o", obj], CSRC)
            return obj
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
            objs = list(pool.map(compile_one, HIP_SOURCES))
        _run([hipcc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", LIB + ".tmp"] + objs,
             CSRC)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_host(cc="gcc"):
    """libeif_host.so (no GPU dependency) and eif_gpu_main (needs libeif_gpu.so)."""
    cmd = [cc] + HOST_CFLAGS + ["-fPIC", "-shared", "-o", HOSTLIB + ".tmp"] + HOST_SOURCES + ["-lm"]
    _run(cmd, HOST)
    os.replace(HOSTLIB + ".tmp", HOSTLIB)
    cmd = [cc] + HOST_CFLAGS + ["-I", os.path.join(ROOT, "include"), "-o", EXE + ".tmp",
                                "main.c"] + HOST_SOURCES + [
        "-L", HERE, "-leif_gpu", "-Wl,-rpath,$ORIGIN", "-lm"]
    _run(cmd, HOST)
    os.replace(EXE + ".tmp", EXE)
    return EXE


def build_all():
    build_gpu()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
