"""Seeded workloads of the benchmark configurations and the batched / served
measurements built on them.

  advection_c1()    2D advection p=1, 32x32 jittered squares of [-1,1]^2, to t=2
  vortex_c2()       isentropic vortex p=2, 128x128 squares of [0,10]^2, to t=2
  vortex3_c3()      3 superposed vortices, p=4, 512x512 squares (524,288 triangles)
  ragged_batch_c4() 64 solves, log-uniform 500..200k triangles, p in 1..4,
                    half advection (Gaussian pulses), half Euler vortices
  serving_c5()      Poisson request stream, log-uniform 1k..500k triangles,
                    p in 1..3, t_end in [0.5, 2]
  serving_small()   the same stream with 1k..20k triangles
"""
import math
import time

import numpy as np

from . import (DG_FP32, DG_FP64, Batch, Mesh, Server, Solver, advection_gaussian, euler_vortex, n_modes,
               request)


def nx_for(triangles):
    return max(2, int(round(math.sqrt(triangles / 2.0))))


def vortices(seed, n=3, L=10.0):
    rng = np.random.default_rng(seed)
    return tuple((float(rng.uniform(2, 5)), float(rng.uniform(0, L)), float(rng.uniform(0, L))) for _ in range(n))


def advection_c1(seed=1):
    m = Mesh.periodic_jittered(32, 32, -1.0, 1.0, -1.0, 1.0, 0.1, seed)
    return m, advection_gaussian(a=(1.0, 1.0), center=(0.0, 0.0), k=40.0), 1, 2.0


def vortex_c2(seed=1, n=128):
    m = Mesh.periodic_jittered(n, n, 0.0, 10.0, 0.0, 10.0, 0.1, seed)
    return m, euler_vortex(((5.0, 5.0, 5.0),)), 2, 2.0


def vortex3_c3(seed=2024, n=512):
    m = Mesh.periodic_jittered(n, n, 0.0, 10.0, 0.0, 10.0, 0.1, seed)
    return m, euler_vortex(vortices(seed)), 4


def ragged_batch_c4(seed=7, n=64, kmin=500, kmax=200_000):
    """(mesh, problem, p) for each solve of the ragged batch."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        K = math.exp(rng.uniform(math.log(kmin), math.log(kmax)))
        nx = nx_for(K)
        p = int(rng.integers(1, 5))
        m = Mesh.periodic_jittered(nx, nx, 0.0, 10.0, 0.0, 10.0, 0.1, int(rng.integers(1, 2**31)))
        if i % 2 == 0:
            ang = rng.uniform(0, 2 * math.pi)
            speed = rng.uniform(0.1, 1.0)
            prob = advection_gaussian(a=(speed * math.cos(ang), speed * math.sin(ang)),
                                      center=(float(rng.uniform(0, 10)), float(rng.uniform(0, 10))), k=1.0)
        else:
            prob = euler_vortex(((float(rng.uniform(2, 5)), float(rng.uniform(0, 10)), float(rng.uniform(0, 10))),))
        out.append((m, prob, p))
    return out


def serving_c5(seed=11, n=32, rate=40.0, kmin=1000, kmax=500_000):
    """Request stream: Poisson arrivals at `rate` per second (offsets from 0)."""
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        K = math.exp(rng.uniform(math.log(kmin), math.log(kmax)))
        nx = nx_for(K)
        p = int(rng.integers(1, 4))
        te = float(rng.uniform(0.5, 2.0))
        prob = euler_vortex(((float(rng.uniform(2, 5)), float(rng.uniform(0, 10)), float(rng.uniform(0, 10))),))
        reqs.append(request(nx, p, te, prob, mesh_seed=int(rng.integers(1, 2**31)), arrival_s=t))
    return reqs


def serving_small(seed=13, n=128, rate=1000.0, kmin=1000, kmax=20_000):
    """The serving stream restricted to small solves (1k-20k triangles):
    the regime where one-at-a-time solving is bound by launch latency."""
    return serving_c5(seed=seed, n=n, rate=rate, kmin=kmin, kmax=kmax)


def dof_stage_per_step(mesh, prob, p):
    return mesh.num_elem * n_modes(p) * prob.physics * 4


def kernel_bytes_per_elem(N, n_p, s, kernel):
    """Bytes the kernels of this build must move per element and stage
    (each input read once, each output written once; halo re-reads not
    counted).  Every kernel runs RK4 in stage states: state passes
    3.0*N*n_p*s (stage 1 reads u_n and writes u2; stage 2 reads u2 and u_n
    and writes u3 and W = (2 u_n + (u2 - u_n) + dt k2)/3; stage 3 reads u3
    and u_n and writes u4; stage 4 reads u4 and W and writes u_n+1: 2, 4,
    3, 3 arrays).  Matrix-core path (kernel 2/3): the packed geometry, per
    face one record of 4 values (nx, ny, len, neighbour word), per element a
    map of 4 values (xr, yr, xs, ys) and 1/J, 17*s bytes (68 in FP32).
    Lane-per-element path (kernel 1): geometry 14*s (xr, yr, xs, ys, 1/J; per
    face nx, ny, len) and 3 32-bit neighbour words."""
    if kernel >= 2:
        return 3.0 * N * n_p * s + (3 * 4 + 4 + 1) * s
    return 3.0 * N * n_p * s + 14 * s + 3 * 4


def bytes_per_step(mesh, prob, p, precision, kernel):
    """HBM bytes of one RK4 step (4 stages) of one solve on kernel `kernel`."""
    s = 4 if precision == DG_FP32 else 8
    return 4 * mesh.num_elem * kernel_bytes_per_elem(prob.physics, n_modes(p), s, kernel)


def _ceiling(runs, elapsed, detail):
    """Bounds of any batching of `runs` = [(p, DOF*stage, HBM bytes, seconds
    alone)], which took `elapsed` one at a time:
      ideal_s: the work at the best rate a single solve of each order p
        reached (that of the largest); small solves alone are bound by launch
        latency, which a batch amortises, so a batch can beat this figure;
      hbm_bytes: the bytes the kernels must move; divided by a measured copy
        rate this is the least time any schedule of these kernels takes."""
    best = {}
    for p, work, _, ti in runs:
        best[p] = max(best.get(p, 0.0), work / ti)
    ideal = sum(work / best[p] for p, work, _, _ in runs)
    detail.update(best_rate_by_p={str(p): best[p] for p in sorted(best)}, ideal_s=ideal,
                  ideal_ratio=elapsed / ideal, hbm_bytes=float(sum(b for _, _, b, _ in runs)),
                  work_dof_stage=float(sum(w for _, w, _, _ in runs)),
                  alone_s=float(sum(t for _, _, _, t in runs)))


def run_batch(cases, steps, precision=DG_FP32, use_graph=True, warmup=2, launch_mode=None):
    """All cases in one batch, `steps` RK4 steps each; returns (seconds,
    DOF*stage, batch).  launch_mode: DG_BATCH_ONE_LAUNCH or
    DG_BATCH_PER_CLASS (None: the batch's default for the precision)."""
    b = Batch(precision)
    if launch_mode is not None:
        b.set_launch_mode(launch_mode)
    for m, prob, p in cases:
        b.add(m, prob, p, 1e30)
    b.project_ic()
    b.run_steps(warmup, 0.3, use_graph)
    t0 = time.perf_counter()
    b.run_steps(steps, 0.3, use_graph)
    el = time.perf_counter() - t0
    work = sum(dof_stage_per_step(m, prob, p) for m, prob, p in cases) * steps
    return el, work, b


def run_sequential(cases, steps, precision=DG_FP32, use_graph=True, warmup=2, detail=None):
    """The same cases one after another on the single-solve path.  With a
    dict `detail`, also each solve's time alone and the bounds of any
    batching of these cases (_ceiling)."""
    solvers = []
    for m, prob, p in cases:
        s = Solver(m, prob, p, precision)
        s.project_ic()
        s.run_steps(warmup, 0.3, use_graph=use_graph)
        solvers.append(s)
    times = []
    t0 = time.perf_counter()
    for s in solvers:
        ti = time.perf_counter()
        s.run_steps(steps, 0.3, use_graph=use_graph)  # returns once the steps are done
        times.append(time.perf_counter() - ti)
    el = time.perf_counter() - t0
    works = [dof_stage_per_step(m, prob, p) * steps for m, prob, p in cases]
    if detail is not None:
        runs = [(p, w, bytes_per_step(m, prob, p, precision, s.kernel()) * steps, ti)
                for (m, prob, p), w, s, ti in zip(cases, works, solvers, times)]
        _ceiling(runs, el, detail)
    return el, sum(works), solvers


def _latency_stats(lat):
    lat = np.sort(np.asarray(lat))
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()))


def run_served(reqs, precision=DG_FP32, max_slots=64, hbm_budget=64 << 30, steps_per_round=8, keep=()):
    """Serve the request stream; returns metrics and the kept final states."""
    sv = Server(precision, max_slots=max_slots, hbm_budget=hbm_budget, steps_per_round=steps_per_round, cfl=0.3)
    sv.keep_states(bool(keep))
    t0 = sv.now() + 0.05
    ids = []
    for r in reqs:
        rr = request(r.nx, r.p, r.t_end, r.prob, mesh_seed=r.mesh_seed, arrival_s=t0 + r.arrival_s)
        ids.append(sv.submit(rr))
    sv.drain()
    res = {x.id: x for x in sv.poll(len(reqs) + 16)}
    lat = [res[i].finish_s - res[i].arrival_s for i in ids]
    makespan = max(res[i].finish_s for i in ids) - t0
    states = {k: sv.get_state(ids[k], (4, n_modes(reqs[k].p), 2 * reqs[k].nx * reqs[k].ny)) for k in keep}
    bad = sum(1 for i in ids if res[i].status != 0)
    drift = max(abs(res[i].mass[0] - res[i].mass0[0]) / abs(res[i].mass0[0]) for i in ids)
    sv.close()
    out = dict(solves=len(reqs), failed=bad, makespan_s=makespan, solves_per_s=len(reqs) / makespan,
               max_rel_mass_drift=drift)
    out.update({"latency_" + k: v for k, v in _latency_stats(lat).items()})
    return out, states


def run_plain(reqs, precision=DG_FP32, keep=()):
    """One request at a time on the single-solve path, in arrival order:
    build the mesh and solver, project, integrate to t_end, audit, free.
    Also each request's integration time alone and the bounds of any
    batching of this stream (_ceiling)."""
    t0 = time.perf_counter() + 0.05
    lat, states, runs = [], {}, []
    for k, r in enumerate(reqs):
        now = time.perf_counter()
        if now < t0 + r.arrival_s:
            time.sleep(t0 + r.arrival_s - now)
        m = Mesh.periodic_jittered(r.nx, r.ny, r.x0, r.x1, r.y0, r.y1, r.jitter, r.mesh_seed)
        s = Solver(m, r.prob, r.p, precision)
        s.project_ic()
        s.sync()
        ti = time.perf_counter()
        n, _ = s.integrate(r.t_end, 0.3)
        ti = time.perf_counter() - ti
        s.audit()
        if k in keep:
            states[k] = s.get_state()
        runs.append((r.p, dof_stage_per_step(m, r.prob, r.p) * n,
                     bytes_per_step(m, r.prob, r.p, precision, s.kernel()) * n, ti))
        s.close()
        lat.append(time.perf_counter() - (t0 + r.arrival_s))
    makespan = time.perf_counter() - t0
    out = dict(solves=len(reqs), makespan_s=makespan, solves_per_s=len(reqs) / makespan)
    out.update({"latency_" + k: v for k, v in _latency_stats(lat).items()})
    _ceiling(runs, makespan, out)
    return out, states
