"""The reference-side C binding (dgcuda_amd/binding: run_dgcuda and the
supersonic vortex's problem unit) compiles against include/dgmi.h and links
libdgmi.so with a plain C compiler.  On the CPU it runs up to the solver,
which reports that no GPU is visible; on the GPU it runs the reference's
supersonic-vortex case on sv1 and writes the density view."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIND = os.path.join(ROOT, "dgcuda_amd", "binding")
LIBDIR = os.path.join(ROOT, "dgcuda_amd")
SV1 = os.path.join(ROOT, "tests", "golden", "sv1.pmsh")


def build(tmp_path):
    exe = str(tmp_path / "supersonic")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(BIND, "run_dgcuda.c"), os.path.join(BIND, "supersonic.c"), "-L", LIBDIR, "-ldgmi",
                    "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_binding_compiles_and_reports_missing_gpu(tmp_path, dg):
    exe = build(tmp_path)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "usage" in r.stderr
    if dg.device_count() > 0:
        pytest.skip("a GPU is visible: the GPU test runs the case")
    r = subprocess.run([exe, "-n", "2", "-T", "0.01", SV1], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, cwd=tmp_path)
    assert r.returncode == 1 and "no GPU" in r.stderr


def test_integration_notes_show_the_compiled_binding():
    """The run_dgcuda body the integration notes print is the file this test
    module compiles (skipped where the notes are not shipped)."""
    notes = os.path.join(ROOT, "INTEGRATION.md")
    if not os.path.exists(notes):
        pytest.skip("INTEGRATION.md not present")
    with open(notes) as f:
        text = f.read()
    m = re.search(r"```c\n(/\* run_dgcuda\.c.*?)```", text, re.S)
    assert m, "no run_dgcuda.c block in INTEGRATION.md"
    with open(os.path.join(BIND, "run_dgcuda.c")) as f:
        assert m.group(1).strip() == f.read().strip()


@pytest.mark.gpu
def test_binding_runs_the_supersonic_vortex(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "-n", "2", "-T", "0.01", SV1], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.match(r"(\d+) steps to t = (\S+)", r.stdout)
    assert m and int(m.group(1)) > 0 and float(m.group(2)) == 0.01
    text = open(tmp_path / "U0.msh").read()
    assert text.startswith('View "U0 " {') and text.count("ST (") == 166
