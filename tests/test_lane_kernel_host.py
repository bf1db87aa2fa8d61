"""Host build of the lane-per-element stage body with its boundary path at
every order, run under AddressSanitizer over every element (padding
included) and every stage kind, with buffers of exactly the device sizes
(tests/host_check/lane_bounds.cpp).  CPU only.

It checks the global-memory indexing of valu_process, in particular of the
p = 4 FP64 instance with boundary sides, whose first GPU run faulted (memory
aperture violation): every access stays inside its buffer on the shipped
boundary meshes, so the fault's cause is not in that indexing, and the GPU
keeps the instance without the path (the product caps the path at p = 3).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dgcuda_amd", "csrc")
HERE = os.path.join(ROOT, "tests", "host_check")
CLANG = "/opt/rocm/llvm/bin/clang++"
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("clang++ with AddressSanitizer not available")
    exe = str(tmp_path_factory.mktemp("host_check") / "lane_bounds")
    cmd = [CLANG, "-std=c++17", "-O0", "-fsanitize=address", "-DDG_GLOBAL=", "-DDGMI_VALU_BOUNDARY_MAX_P=5",
           "-I", HERE, "-I", CSRC, os.path.join(HERE, "lane_bounds.cpp")] + \
          [os.path.join(CSRC, f) for f in ("dg_geometry.cpp", "dg_mesh.cpp", "dg_tables.cpp")] + ["-o", exe]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


CASES = [("sv1refined.pmsh", N, p, prec) for N in (1, 4) for p in (1, 2, 3, 4, 5) for prec in (0, 1)]
CASES += [("sv1refined1.pmsh", 4, 4, 0), (os.path.join("gmsh", "cylinder1.pmsh"), 4, 4, 0)]


@pytest.mark.parametrize("mesh,N,p,prec", CASES)
def test_lane_body_accesses_in_bounds(harness, mesh, N, p, prec):
    r = subprocess.run([harness, os.path.join(GOLD, mesh), str(N), str(p), str(prec)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "every access in bounds" in r.stdout


@pytest.mark.parametrize("N,p,prec", [(4, 1, 1), (4, 2, 1), (4, 2, 0), (1, 2, 1)])
def test_lane_body_accesses_in_bounds_periodic(harness, tmp_path, N, p, prec):
    """The variant without boundary code (periodic meshes), whose neighbour
    words and first neighbour's modes are loaded ahead of the volume
    integral, on a jittered periodic mesh written as .pmsh."""
    import dgcuda_amd as dg
    path = str(tmp_path / "periodic.pmsh")
    dg.Mesh.periodic_jittered(12, 10, 0.0, 10.0, 0.0, 10.0, 0.1, 3).write_pmsh(path)
    r = subprocess.run([harness, path, str(N), str(p), str(prec)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "every access in bounds" in r.stdout and "boundary faces 0" in r.stdout
