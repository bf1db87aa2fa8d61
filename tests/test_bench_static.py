"""bench.py and the workload module run only on a GPU; on the CPU this checks
that every name their functions read is defined (a local, a module-level
name or a builtin), and the pure helpers of the bench line."""
import ast
import builtins
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["bench.py", "dgcuda_amd/workloads.py", "__graft_entry__.py", "oracle/cpu_baseline.py"]


def _module_names(tree):
    names = set(dir(builtins))
    for node in tree.body:
        if isinstance(node, (ast.FunctionDef, ast.ClassDef)):
            names.add(node.name)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            names.update((a.asname or a.name).split(".")[0] for a in node.names)
        else:
            for x in ast.walk(node):
                if isinstance(x, ast.Name) and isinstance(x.ctx, ast.Store):
                    names.add(x.id)
    return names


def _undefined(tree):
    glob = _module_names(tree)
    out = []
    for f in (n for n in ast.walk(tree) if isinstance(n, (ast.FunctionDef, ast.Lambda))):
        local = set()
        for x in ast.walk(f):
            if isinstance(x, ast.Name) and isinstance(x.ctx, (ast.Store, ast.Del)):
                local.add(x.id)
            elif isinstance(x, ast.arg):
                local.add(x.arg)
            elif isinstance(x, (ast.Import, ast.ImportFrom)):
                local.update((a.asname or a.name).split(".")[0] for a in x.names)
            elif isinstance(x, ast.FunctionDef):
                local.add(x.name)
            elif isinstance(x, ast.ExceptHandler) and x.name:
                local.add(x.name)
        for x in ast.walk(f):
            if isinstance(x, ast.Name) and isinstance(x.ctx, ast.Load) and x.id not in local | glob:
                out.append((getattr(f, "name", "<lambda>"), x.id, x.lineno))
    return out


@pytest.mark.parametrize("rel", FILES)
def test_every_name_is_defined(rel):
    with open(os.path.join(ROOT, rel)) as fh:
        tree = ast.parse(fh.read())
    assert _undefined(tree) == []


def test_checker_finds_a_missing_import():
    tree = ast.parse("def f():\n    return W.g()\n\ndef h():\n    from x import y as W\n    return W\n")
    assert _undefined(tree) == [("f", "W", 2)]


def test_bounds_of_batching():
    import bench
    from dgcuda_amd import workloads as W

    d = {}
    # (p, DOF*stage, HBM bytes, seconds alone): two solves, 0.2 s one at a time
    W._ceiling([(1, 1e9, 2e10, 0.1), (4, 5e9, 6e10, 0.05)], 0.2, d)
    assert d["ideal_s"] == pytest.approx(0.15) and d["hbm_bytes"] == 8e10
    b = bench.bounds(d, 0.2, 5000.0)
    assert b["hbm_floor_s"] == pytest.approx(8e10 / 5e12)
    assert b["ratio_at_hbm_floor"] == pytest.approx(0.2 / 0.016, rel=1e-3)
    assert b["ratio_at_single_rates"] == pytest.approx(0.2 / 0.15, rel=1e-3)


def test_kernel_byte_model():
    from dgcuda_amd import workloads as W

    # Euler p = 4 FP32 on the matrix-core path: 12 state arrays per step
    # (3 per stage) of 60 values and 68 B of packed geometry per element
    assert W.kernel_bytes_per_elem(4, 15, 4, 2) == 3 * 60 * 4 + 68
    # lane path: the same 3 state passes, 14 geometry values, 3 neighbour words
    assert W.kernel_bytes_per_elem(4, 3, 4, 1) == 3 * 12 * 4 + 14 * 4 + 12
