"""Host build of the batch's work-list partition (dgcuda_amd/csrc/dg_worklist.h,
tests/host_check/worklist_check.cpp): walked as the persistent batch
kernels walk it, every (solve, tile) item is taken exactly once, in list
order, from 8 XCD chunks of about equal cost.  CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dgcuda_amd", "csrc")
HERE = os.path.join(ROOT, "tests", "host_check")


def test_worklist_layout(tmp_path):
    exe = str(tmp_path / "worklist_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address", "-I", HERE, "-I", CSRC,
                    os.path.join(HERE, "worklist_check.cpp"), "-o", exe],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "worklist partition ok" in r.stdout
