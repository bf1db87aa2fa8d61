#!/usr/bin/env python3
"""Euler p=4 DOF.stage updates/s on one MI355X (the 3-vortex, 524,288-triangle
periodic workload), with the stage kernel's HBM roofline and the FP64 CPU
oracle's throughput beside it.

  python bench.py [--gpus 1] [--steps 200] [--warmup 10]

Prints one JSON line.  Beside the headline workload it carries C1 and C2
run to their end times ("configs": steps, time to solution, errors), the
ragged batch ("batch": 64 solves in one launch per RK stage vs one at a
time) and the serving loop ("serving": served vs plain solves/s and
p50/p99 latency on seeded Poisson request streams).  A step is one classical RK4 step (4 stages) of the
whole mesh with the CFL time step computed on the device; the timed region
replays the captured step graph `--steps` times with the state resident in
HBM.
"""
import argparse
import hashlib
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, published


def vortices(seed, n=3, L=10.0):
    rng = np.random.default_rng(seed)
    return tuple((float(rng.uniform(2, 5)), float(rng.uniform(0, L)), float(rng.uniform(0, L))) for _ in range(n))


def rel_l2(a, b):
    """Relative distance of two modal states: sqrt(sum |a - b|^2) /
    sqrt(sum |b|^2) over every coefficient (Parseval on the orthonormal
    basis, elements weighted alike)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(b ** 2)))


def algorithmic_bytes_per_elem(N, n_p, s):
    """Algorithmic bytes of one RK4 stage per element, the reference
    algorithm's byte model: state and RK registers 4.25*N*n_p*s (classic RK4
    with a fused stage update: each stage reads its input once; stage 1
    writes acc and u2; stages 2-3 read u_n and acc and write acc and the
    next input; stage 4 reads u_n and acc and writes u_n+1, i.e. (1+2, 1+4,
    1+4, 1+3)/4 = 4.25 streams) plus 110 B of geometry (FP64 metrics: 5x8 B
    per element, 3x4 B element-to-face indices, 1.5 faces x (3x8 B normal
    and length + 3x4 B connectivity), CFL cell means N*s/4).  Face-flux
    buffers are not counted.  C3 FP32: 18.83 B per DOF*stage."""
    return 4.25 * N * n_p * s + 110.0


def start_cpu_baseline(c3_steps):
    """The FP64 serial oracle on one pinned host core (oracle/cpu_baseline.py)
    in a child process, started before this process touches the GPU and run
    beside the GPU measurements."""
    return subprocess.Popen([sys.executable, "-m", "oracle.cpu_baseline", "--c3-steps", str(c3_steps)], cwd=ROOT,
                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def finish_cpu_baseline(proc):
    out, err = proc.communicate(timeout=1800)
    if proc.returncode != 0:
        raise RuntimeError("CPU baseline failed (exit %d): %s" % (proc.returncode, err[-2000:]))
    d = json.loads(out.strip().splitlines()[-1])
    c3, c1, pl = d["c3"], d["c1"], d["c3_plain"]
    sample = ("FP64 serial oracle, 1 thread pinned to core %d of %d (%s): %d full RK4 steps of this workload "
              "(%d triangles, p=4) in %.1f s, 200 steps extrapolated to %.0f s"
              % (d["pinned_core"], d["host_logical_cores"], d["cpu_model"], c3["steps"], c3["triangles"],
                 c3["seconds"], c3["extrapolated_200_steps_s"]))
    return dict(value=c3["value"], unit="DOF*stage/s", cores=1, kind="port", sample=sample,
                cpu_model=d["cpu_model"], host_logical_cores=d["host_logical_cores"], pinned_core=d["pinned_core"],
                extrapolated_200_steps_s=c3["extrapolated_200_steps_s"],
                advection_p1_2048_full_run=dict(value=c1["value"], steps=c1["steps"], seconds=c1["seconds"]),
                reference_structured=dict(value=pl["value"], seconds=pl["seconds"],
                                          sample="the same oracle without hoisting (traces and fluxes recomputed "
                                                 "per mode, as the reference's loops do), %d RK4 steps on a "
                                                 "64x64-cell patch (%d triangles) of the workload"
                                                 % (pl["steps"], pl["triangles"])))


# stage-kernel sources whose change makes a recorded traffic profile stale
STAGE_SOURCES = ["dg_mfma_body.h", "dg_mfma.hip", "dg_device.h", "dg_kernels.h"]


def stage_sources_sha256():
    h = hashlib.sha256()
    for name in STAGE_SOURCES:
        with open(os.path.join(ROOT, "dgcuda_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def load_traffic(args, kernel):
    """Measured HBM bytes per stage launch (rocprofv3 FETCH_SIZE / WRITE_SIZE,
    calibrated on a copy kernel), committed under profiles/ for the default
    workload; None for any other configuration."""
    if (args.p, args.n, args.precision, args.seed) != (4, 512, "fp32", 2024):
        return None
    path = os.path.join(ROOT, "profiles", "traffic_euler_p4_524k.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("kernel") != kernel:
        return None
    if d.get("sources_sha256") != stage_sources_sha256():
        return "stale"  # profiled on other stage-kernel sources
    return d.get("bytes_per_launch")


def bounds(detail, elapsed, copy_gbs):
    """The bounds of batching a workload that took `elapsed` s one solve at a
    time (workloads._ceiling): at the single-solve rates, and at the copy
    kernel's HBM rate over the bytes the kernels must move."""
    floor = detail["hbm_bytes"] / (copy_gbs * 1e9)
    return dict(best_single_rate_by_p={k: float("%.4g" % v) for k, v in detail["best_rate_by_p"].items()},
                ideal_at_single_rates_s=float("%.4g" % detail["ideal_s"]),
                ratio_at_single_rates=round(elapsed / detail["ideal_s"], 3),
                hbm_bytes=float("%.4g" % detail["hbm_bytes"]),
                hbm_floor_s=float("%.4g" % floor), ratio_at_hbm_floor=round(elapsed / floor, 3))


def batch_summary(copy_gbs, steps=20, steps64=4):
    """Ragged batch: 64 solves (500-200k triangles, p 1..4, half advection
    and half Euler vortices) in each launch form of the batch, against
    the same solves one after another on the single-solve path, and the
    bounds of any batching of them."""
    import dgcuda_amd as dg
    from dgcuda_amd import workloads as W

    cases = W.ragged_batch_c4()
    out = dict(workload="ragged_batch_64_solves_500_to_200k_p1to4", solves=len(cases),
               triangles=sum(m.num_elem for m, _, _ in cases))
    forms = ((dg.DG_BATCH_ONE_LAUNCH, "one_launch"), (dg.DG_BATCH_PER_CLASS, "per_class"))
    for prec, name, n in ((dg.DG_FP32, "fp32", steps), (dg.DG_FP64, "fp64", steps64)):
        el = {}
        for mode, fname in forms:
            el[mode], work, b = W.run_batch(cases, n, prec, launch_mode=mode)
            b.close()
        detail = {}
        el2, _, ss = W.run_sequential(cases, n, prec, detail=detail)
        for x in ss:
            x.close()
        dflt = dg.BATCH_DEFAULT_LAUNCH_MODE[prec]
        out[name] = dict(steps=n, launch_form=dict(forms)[dflt],
                         batched_dof_stage_per_s=float("%.4g" % (work / el[dflt])),
                         one_at_a_time_dof_stage_per_s=float("%.4g" % (work / el2)),
                         speedup=round(el2 / el[dflt], 3),
                         forms={fname: dict(dof_stage_per_s=float("%.4g" % (work / el[mode])),
                                            speedup=round(el2 / el[mode], 3)) for mode, fname in forms},
                         bounds=bounds(detail, el2, copy_gbs))
    out["kernel_choice"] = ("fixed per order: lane items (one element per lane) for p <= 2, matrix-core tiles "
                            "for p >= 3 (the autotuner's choice for single FP32 solves at large sizes; for FP64 Euler "
                            "p = 2 it picks tiles, which batched measured no faster)")
    out["launches"] = ("forms: one_launch = one persistent launch per RK stage carrying every kernel class; "
                       "per_class = one per stage and kernel class (Euler/advection x tiles/lane items), each at "
                       "its own occupancy; speedup is the batch's default form (launch_form) in each precision")
    return out


def configs_summary():
    """C1 and C2 on the GPU, to their end time: steps, time to solution and
    the errors the configurations report.  C1 (advection p = 1, FP64): L2
    error against the initial pulse after one period.  C2 (vortex p = 2):
    FP64 and FP32 runs, FP32 against FP64, each against the exact vortex
    carried to (7, 7), and the mass drift of each."""
    import dgcuda_amd as dg
    from dgcuda_amd import workloads as W

    out = {}
    m, prob, p, t_end = W.advection_c1()
    s = dg.Solver(m, prob, p, dg.DG_FP64)
    s.project_ic()
    u0 = s.get_state()
    s.sync()
    t0 = time.perf_counter()
    n, _ = s.integrate(t_end, 0.3)
    el = time.perf_counter() - t0
    out["c1_advection_p1_2048_fp64"] = dict(steps=n, seconds=round(el, 4),
                                            rel_l2_vs_initial=float("%.3e" % rel_l2(s.get_state(), u0)))
    s.close()
    m, prob, p, t_end = W.vortex_c2()
    exact = dg.Solver(m, dg.euler_vortex(((5.0, 7.0, 7.0),)), p, dg.DG_FP64)
    exact.project_ic()
    ue = exact.get_state()
    exact.close()
    res = {}
    for name, prec in (("fp64", dg.DG_FP64), ("fp32", dg.DG_FP32)):
        s = dg.Solver(m, prob, p, prec)
        s.project_ic()
        m0 = np.array(s.audit()["mass"])
        s.sync()
        t0 = time.perf_counter()
        n, _ = s.integrate(t_end, 0.3)
        el = time.perf_counter() - t0
        st = s.get_state()
        drift = float(np.max(np.abs(np.array(s.audit()["mass"]) - m0) / np.abs(m0)))
        res[name] = (st, dict(steps=n, seconds=round(el, 4), rel_l2_vs_exact=float("%.3e" % rel_l2(st, ue)),
                              max_rel_mass_drift=float("%.2e" % drift), kernel=s.kernel()))
        s.close()
    c2 = {k: v[1] for k, v in res.items()}
    c2["rel_l2_fp32_vs_fp64"] = float("%.3e" % rel_l2(res["fp32"][0], res["fp64"][0]))
    out["c2_vortex_p2_32768"] = c2
    return out


def serving_summary(copy_gbs):
    """Served (continuous batching) against plain one-at-a-time solves on the
    same seeded Poisson request streams: the 1k-500k-triangle mix and the
    small-solve mix, and the bounds of any batching of them."""
    from dgcuda_amd import workloads as W

    out = {}
    for name, reqs in (("stream_1k_to_500k_p1to3_32req_40per_s", W.serving_c5()),
                       ("stream_1k_to_20k_p1to3_128req_1000per_s", W.serving_small())):
        served, _ = W.run_served(reqs)
        plain, _ = W.run_plain(reqs)
        def r(d):
            return {k: (float("%.4g" % v) if isinstance(v, float) else r(v) if isinstance(v, dict) else v)
                    for k, v in d.items()}
        b = bounds(plain, plain["makespan_s"], copy_gbs)
        for k in ("best_rate_by_p", "ideal_s", "ideal_ratio", "hbm_bytes"):
            plain.pop(k)
        out[name] = dict(served=r(served), plain=r(plain), bounds=b,
                         solves_per_s_ratio=round(served["solves_per_s"] / plain["solves_per_s"], 3))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="cells per side (2*n*n triangles)")
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 VALU, 2 matrix cores")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-c3-steps", type=int, default=2, help="full RK4 steps of the CPU baseline sample")
    ap.add_argument("--prewarm-s", type=float, default=2.0,
                    help="seconds of graph-replayed steps before the warm-up (lets the clocks settle)")
    ap.add_argument("--accuracy-steps", type=int, default=200,
                    help="steps of the FP32 vs FP64 accuracy check (separate solves from the initial state)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the ragged-batch measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurements")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1 / C2 runs")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many steps (for rocprof) and print nothing else")
    ap.add_argument("--profile-graph", action="store_true",
                    help="with --profile-steps: replay the captured step graph, as the timed run does")
    ap.add_argument("--profile-batch", default="", choices=["", "fp32", "fp64"],
                    help="with --profile-steps: run only the ragged batch (C4) in this precision")
    ap.add_argument("--profile-per-class", action="store_true",
                    help="with --profile-batch: one launch per stage and kernel class")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    cpu_proc = None
    if not args.no_cpu and not args.profile_steps:
        cpu_proc = start_cpu_baseline(args.cpu_c3_steps)

    import dgcuda_amd as dg
    from dgcuda_amd import workloads as W

    if args.profile_steps and args.profile_batch:
        bprec = dg.DG_FP32 if args.profile_batch == "fp32" else dg.DG_FP64
        mode = dg.DG_BATCH_PER_CLASS if args.profile_per_class else dg.DG_BATCH_ONE_LAUNCH
        el, work, b = W.run_batch(W.ragged_batch_c4(), args.profile_steps, bprec, launch_mode=mode)
        b.close()
        print("profiled %d batch steps in %.4f s" % (args.profile_steps, el), flush=True)
        return

    prec = dg.DG_FP32 if args.precision == "fp32" else dg.DG_FP64
    s_bytes = 4 if prec == dg.DG_FP32 else 8
    t_setup = time.perf_counter()
    mesh = dg.Mesh.periodic_jittered(args.n, args.n, 0.0, 10.0, 0.0, 10.0, 0.1, args.seed)
    vs = vortices(args.seed)
    prob = dg.euler_vortex(vs)
    s = dg.Solver(mesh, prob, args.p, prec)
    if args.kernel:
        s.set_kernel(args.kernel)
    s.project_ic()
    s.sync()
    setup_s = time.perf_counter() - t_setup
    if args.profile_steps:
        s.run_steps(args.profile_steps, 0.3, use_graph=args.profile_graph)
        s.sync()
        print("profiled %d steps" % args.profile_steps, flush=True)
        return

    K, n_p, N = s.num_elem, s.n_p, s.N
    dof = K * n_p * N
    # pre-warm: replayed steps until the clocks have settled, then the
    # requested warm-up; neither is timed
    t_pw, n_pw = time.perf_counter(), 0
    while time.perf_counter() - t_pw < args.prewarm_s:
        s.run_steps(20, 0.3, use_graph=True)
        s.sync()
        n_pw += 20
    prewarm = dict(seconds=round(time.perf_counter() - t_pw, 3), steps=n_pw)
    s.run_steps(args.warmup, 0.3, use_graph=True)
    s.sync()
    # timed region: the captured step graph replayed `steps` times, with HIP
    # events on the solver's stream around the replays (for the stage kernel's
    # live duration) inside the wall-clock bracket (for `value`)
    t0 = time.perf_counter()
    tg = s.time_step_graph(args.steps, 0.3)
    s.sync()
    el = time.perf_counter() - t0
    value = dof * 4 * args.steps / el
    t_reached = s.audit()["t"]

    # dominant kernel: the fused stage kernel.  Its live duration comes from
    # the timed region's events: ms per replayed step, less the dt kernel
    # (timed per launch), over the 4 stage launches of a step.  Event pairs
    # around each directly launched stage (stage_ms_direct) separate the
    # kernels and read longer.
    tk = s.time_stage_kernels(20, 0.3)
    stage_ms = (tg["step_ms"] - tk["dt_ms"]) / 4
    bpe = algorithmic_bytes_per_elem(N, n_p, s_bytes)
    bytes_launch = bpe * K
    achieved = bytes_launch / (stage_ms * 1e-3) / 1e9
    kbytes_launch = W.kernel_bytes_per_elem(N, n_p, s_bytes, s.kernel()) * K
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), traffic=load_traffic(args, s.kernel()),
                kernel="stage (fused volume + LLF faces + M^-1 + RK update)",
                stage_ms=round(stage_ms, 5), stage_ms_direct=round(tk["stage_ms"], 5),
                graph_step_ms=round(tg["step_ms"], 5), dt_kernel_ms=round(tk["dt_ms"], 5),
                algorithmic_bytes_per_launch=int(bytes_launch), bytes_per_dof_stage=round(bpe / (n_p * N), 3),
                kernel_bytes_per_launch=int(kbytes_launch),
                kernel_bytes_GBs=round(kbytes_launch / (stage_ms * 1e-3) / 1e9, 1),
                kernel_bytes_frac=round(kbytes_launch / (stage_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))

    kernel_kind = s.kernel()
    s.close()  # free the workload before the accuracy, batch and serving runs
    extra = {}
    if not args.no_accuracy and prec == dg.DG_FP32:
        # accuracy of the benchmarked configuration: FP32 for accuracy-steps
        # CFL steps from the projected state, against FP64 to the same time,
        # and an FP64 run with FP32-rounded geometry (the geometry storage's
        # share of the FP32 error)
        a32 = dg.Solver(mesh, prob, args.p, dg.DG_FP32)
        a32.set_kernel(kernel_kind)
        a32.project_ic()
        a32.run_steps(args.accuracy_steps, 0.3)
        t_acc = a32.audit()["t"]
        st32 = a32.get_state()
        a32.close()
        a64 = dg.Solver(mesh, prob, args.p, dg.DG_FP64)
        a64.project_ic()
        n64, _ = a64.integrate(t_acc, 0.3)
        st64 = a64.get_state()
        a64.close()
        g64 = dg.Solver(mesh, prob, args.p, dg.DG_FP64 | dg.DG_GEOM_FP32)
        g64.project_ic()
        g64.integrate(t_acc, 0.3)
        stg = g64.get_state()
        g64.close()
        extra["accuracy"] = dict(
            steps=args.accuracy_steps, t=t_acc, fp64_steps=n64,
            rel_l2_fp32_vs_fp64=float("%.3e" % rel_l2(st32, st64)),
            rel_l2_fp32_geometry_only=float("%.3e" % rel_l2(stg, st64)),
            reference="GPU FP64 run to the same time; the FP64 GPU path matches the FP64 oracle to <=1e-12 in "
                      "tests. rel_l2_fp32_geometry_only: FP64 state and kernels with the geometry rounded "
                      "through FP32, as the FP32 path stores it")
    bw_ms, bw, bw_kernel = dg.bandwidth_probe(1 << 30, 10, with_variant=True)
    configs = None if args.no_configs else configs_summary()
    batch = None if args.no_batch else batch_summary(bw)
    serving = None if args.no_serving else serving_summary(bw)
    cpu = finish_cpu_baseline(cpu_proc) if cpu_proc is not None else None
    traffic = roof["traffic"]
    stale = traffic == "stale"
    if stale:
        roof["traffic"] = traffic = None
    out = {
        "metric": "Euler p=4 DOF*stage updates/s",
        "value": value,
        "unit": "DOF*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == dg.DG_FP32 else "fp64",
        "data": "synthetic: seeded 3-vortex isentropic superposition on a seeded jittered periodic mesh",
        "config": {"workload": "euler_rk4_p%d_3vortex_%dx%d_periodic" % (args.p, args.n, args.n),
                   "triangles": K, "p": args.p, "n_p": n_p, "dof": dof, "cfl": 0.3, "integrator": "rk4",
                   "kernel": kernel_kind,
                   "geometry": "fp64-computed, fp32-stored" if prec == dg.DG_FP32 else "fp64"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "copy_kernel_GBs": round(bw, 1),
        "copy_kernel": bw_kernel,
        # measured HBM traffic per launch over the live duration, against the copy probe
        "traffic_GBs": None if traffic is None else round(traffic / (stage_ms * 1e-3) / 1e9, 1),
        "traffic_frac_of_copy": None if traffic is None else round(traffic / (stage_ms * 1e-3) / 1e9 / bw, 4),
        "traffic_stale": stale,
        "prewarm": prewarm,
        "setup_s": round(setup_s, 3),
        "t_reached": t_reached,
        "host": platform.node(),
    }
    if configs is not None:
        out["configs"] = configs
    if batch is not None:
        out["batch"] = batch
    if serving is not None:
        out["serving"] = serving
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
