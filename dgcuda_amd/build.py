"""Build libdgmi.so (HIP kernels for gfx950 + C++ host code + C ABI) in-tree.

Host-only numerics (mesh, FP64 geometry, basis tables) are compiled with g++
and -ffp-contract=off so each product and sum rounds as written; HIP sources
are compiled with hipcc for gfx950 with packed-FP32 instructions disabled.

    python -m dgcuda_amd.build          # or dgcuda_amd.build.build()
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgmi.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

GXX_SOURCES = ["dg_mesh.cpp", "dg_geometry.cpp", "dg_tables.cpp"]
HIP_SOURCES = ["dg_kernels.hip", "dg_mfma.hip", "dg_solver.hip", "dg_batch.hip", "dg_capi.cpp", "dg_server.cpp"]

GXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def build(verbose=False, out=None, defines=()):
    """Compile every source and link dgcuda_amd/libdgmi.so (always rebuilds).
    `out` and `defines` (extra -D flags for the HIP sources) build a variant
    library elsewhere for kernel comparisons."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    lib = out or LIB
    objdir = OBJ if not out else out + ".objs"
    os.makedirs(objdir, exist_ok=True)
    extra = ["-D" + d for d in defines]

    def comp(src):
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        if src in GXX_SOURCES:
            cmd = ["g++"] + GXX_FLAGS + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC] + HIP_FLAGS + extra + ["-x", "hip", "-c", path, "-o", obj]
        _run(cmd)
        return obj

    srcs = GXX_SOURCES + HIP_SOURCES
    jobs = min(len(srcs), max(1, (os.cpu_count() or 4)), 8)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(comp, srcs))
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
