"""CPU baseline of the benchmark: the FP64 serial oracle timed on one host
core.  Test and measurement infrastructure only (the cpu_baseline leg of
bench.py); the product never imports it.

    python -m oracle.cpu_baseline [--core C] [--c3-steps 2] [--full]

Prints one JSON object.  The process pins itself to one core
(sched_setaffinity) before any work and uses one thread.  Samples:

  c3        the 524,288-triangle p = 4 three-vortex workload, `--c3-steps`
            full RK4 steps (default 2) from its projected initial state;
            the 200-step time is extrapolated from them and labelled so
  c1        the advection configuration (p = 1, 2,048 triangles, t = 0..2),
            the full run
  c3_plain  the reference-structured loops (hoist = 0: traces, fluxes and
            wave speeds recomputed inside every mode loop, as the
            reference's kernels and its CPU port do) on a 64x64-cell patch
            of the c3 workload, 8 RK4 steps
  c2        (--full only, about 15 minutes) the p = 2 vortex configuration
            (32,768 triangles), the full run to t = 2

Rates are DOF*stage/s with DOF = K * n_p * N and 4 stages per RK4 step.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _run(o, u, nsteps=None, t_end=None, cfl=0.3):
    t0 = time.perf_counter()
    if t_end is None:
        u, _ = o.run_steps(u, nsteps, cfl)
        n = nsteps
    else:
        u, n, _ = o.integrate(u, t_end, cfl)
    return u, n, time.perf_counter() - t0


def sample_c3(steps, hoist=1, n=512, seed=2024):
    import dgcuda_amd as dg
    import oracle as orc
    from dgcuda_amd.workloads import vortices

    h = 10.0 / 512
    L = n * h
    vs = tuple((b, x * L / 10.0, y * L / 10.0) for b, x, y in vortices(seed))
    m = dg.Mesh.periodic_jittered(n, n, 0.0, L, 0.0, L, 0.1, seed)
    prob = dg.euler_vortex(vs)
    o = orc.Oracle.from_mesh(m, 4, prob, hoist=hoist)
    u = o.project_ic(orc.ic_array(prob, m.period))
    _, ns, el = _run(o, u, nsteps=steps)
    dof = m.num_elem * o.n_p * o.N
    return dict(triangles=m.num_elem, p=4, dof=dof, steps=ns, seconds=round(el, 3),
                value=dof * 4 * ns / el, hoist=hoist)


def sample_full(cfg):
    import oracle as orc
    from dgcuda_amd import workloads as W

    m, prob, p, t_end = W.advection_c1() if cfg == "c1" else W.vortex_c2()
    o = orc.Oracle.from_mesh(m, p, prob)
    u = o.project_ic(orc.ic_array(prob, m.period))
    _, ns, el = _run(o, u, t_end=t_end)
    dof = m.num_elem * o.n_p * o.N
    return dict(triangles=m.num_elem, p=p, dof=dof, steps=ns, t_end=t_end, seconds=round(el, 3),
                value=dof * 4 * ns / el)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--core", type=int, default=-1, help="host core to pin to (-1: the last allowed one)")
    ap.add_argument("--c3-steps", type=int, default=2)
    ap.add_argument("--full", action="store_true", help="also the full c2 run (about 15 minutes)")
    args = ap.parse_args()
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[-1] if args.core < 0 else args.core
    os.sched_setaffinity(0, {core})
    out = dict(kind="port", unit="DOF*stage/s", threads=1, pinned_core=core, cpu_model=cpu_model(),
               host_logical_cores=os.cpu_count(), cores_allowed=len(allowed),
               build="g++ -O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math (oracle/Makefile)")
    c3 = sample_c3(args.c3_steps)
    c3["extrapolated_200_steps_s"] = round(c3["seconds"] / c3["steps"] * 200, 1)
    out["c3"] = c3
    out["c1"] = sample_full("c1")
    out["c3_plain"] = sample_c3(8, hoist=0, n=64)
    if args.full:
        out["c2"] = sample_full("c2")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
