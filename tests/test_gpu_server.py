"""Serving loop: continuous batching with admission by slots and HBM budget.

Bars: every request's final state in the served mix is bit-identical to the
same solve run alone; budget-limited admission completes every request; a
request whose state goes non-physical fails alone (DG_ERR_NUMERIC) and the
others are unaffected; mass is conserved to round-off per request.
"""
import numpy as np
import pytest

import dgcuda_amd as dg

pytestmark = pytest.mark.gpu


def requests(n, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(n):
        t += float(rng.exponential(0.01))
        nx = int(rng.integers(6, 24))
        p = int(rng.integers(1, 4))
        prob = dg.euler_vortex(((float(rng.uniform(2, 5)), float(rng.uniform(0, 10)), float(rng.uniform(0, 10))),))
        out.append(dg.request(nx, p, float(rng.uniform(0.1, 0.4)), prob, mesh_seed=int(rng.integers(1, 10**6)),
                              arrival_s=t))
    return out


def alone(req, solo_min_elems, prec=dg.DG_FP32):
    """The same request served by itself."""
    sv = dg.Server(prec, max_slots=8, steps_per_round=4, cfl=0.3, solo_min_elems=solo_min_elems,
                   fifo_min_elems=0)
    sv.keep_states(True)
    r = dg.request(req.nx, req.p, req.t_end, req.prob, mesh_seed=req.mesh_seed, arrival_s=0.0)
    i = sv.submit(r)
    sv.drain()
    out = sv.poll(4)
    assert len(out) == 1 and out[0].status == 0
    state = sv.get_state(i, (4, dg.n_modes(req.p), 2 * req.nx * req.ny))
    sv.close()
    return state


@pytest.mark.parametrize("solo_min_elems,fifo_min_elems,prec",
                         [(0, 0, dg.DG_FP32), (400, 0, dg.DG_FP32), (400, 600, dg.DG_FP32), (400, 600, dg.DG_FP64)])
def test_server_results_equal_alone(solo_min_elems, fifo_min_elems, prec):
    """In the mix, every result equals the request served alone; with a
    threshold of 400 elements the larger requests run their own kernels
    inside the step graph while the others share the batch launch; with a
    FIFO threshold of 600 the requests of 600 elements and more advance one
    at a time in arrival order, the others beside them.  In FP64 the batch
    launch is one kernel per class (Euler tiles, lane items, ...)."""
    reqs = requests(10)
    assert any(2 * r.nx * r.ny >= 400 for r in reqs) and any(2 * r.nx * r.ny < 400 for r in reqs)
    assert sum(2 * r.nx * r.ny >= 600 for r in reqs) >= 2
    sv = dg.Server(prec, max_slots=8, steps_per_round=4, cfl=0.3, solo_min_elems=solo_min_elems,
                   fifo_min_elems=fifo_min_elems)
    sv.keep_states(True)
    t0 = sv.now()
    ids = []
    for r in reqs:
        r.arrival_s += t0
        ids.append(sv.submit(r))
    sv.drain()
    res = {r.id: r for r in sv.poll(64)}
    assert len(res) == len(reqs)
    for k in range(len(reqs)):
        r = reqs[k]
        out = res[ids[k]]
        assert out.status == 0 and abs(out.t - r.t_end) == 0.0
        shape = (4, dg.n_modes(r.p), 2 * r.nx * r.ny)
        got = sv.get_state(ids[k], shape)
        assert np.array_equal(got, alone(r, solo_min_elems, prec)), k
    for out in res.values():
        assert out.finish_s >= out.arrival_s
        m0, m1 = np.array(out.mass0), np.array(out.mass)
        assert np.all(np.abs(m1 - m0) <= 2e-6 * np.abs(m0) + 1e-9)
    if fifo_min_elems:
        # the large requests finished in arrival order, each before the next
        # one took a step: admission order is arrival order here
        large = [res[ids[k]] for k in range(len(reqs)) if 2 * reqs[k].nx * reqs[k].ny >= fifo_min_elems]
        assert [x.finish_s for x in large] == sorted(x.finish_s for x in large)


def test_server_solo_turnover():
    """One slot, every solve launched with its own kernels: each solve is
    admitted after the previous one retired and was freed, so the step graph
    must be captured again for each (a new solve can be allocated where a
    deleted one was).  Every result equals the request served alone."""
    reqs = requests(6, seed=11)
    sv = dg.Server(dg.DG_FP32, max_slots=1, steps_per_round=4, cfl=0.3, solo_min_elems=1, prep_threads=1)
    sv.keep_states(True)
    t0 = sv.now()
    ids = []
    for r in reqs:
        r.arrival_s += t0
        ids.append(sv.submit(r))
    sv.drain()
    res = {r.id: r for r in sv.poll(64)}
    assert len(res) == len(reqs)
    for k, r in enumerate(reqs):
        assert res[ids[k]].status == 0
        got = sv.get_state(ids[k], (4, dg.n_modes(r.p), 2 * r.nx * r.ny))
        assert np.array_equal(got, alone(r, 1)), k
    sv.close()


def test_server_budget_admission():
    reqs = requests(8, seed=4)
    one = 40 * 2 * 24 * 24 * 4 * 15 * 4 * 2  # room for about two of the largest at a time
    sv = dg.Server(dg.DG_FP32, max_slots=64, hbm_budget=one, steps_per_round=4, cfl=0.3)
    for r in reqs:
        sv.submit(r)
    sv.drain()
    res = sv.poll(64)
    assert len(res) == len(reqs) and all(r.status == 0 for r in res)


def test_server_bad_request_fails_alone():
    reqs = requests(5, seed=6)
    bad = dg.request(10, 2, 0.2, dg.euler_vortex(((25.0, 5.0, 5.0),)), mesh_seed=1)  # T < 0 at the core
    sv = dg.Server(dg.DG_FP32, max_slots=8, steps_per_round=4, cfl=0.3)
    good_ids = [sv.submit(r) for r in reqs]
    bad_id = sv.submit(bad)
    sv.drain()
    res = {r.id: r for r in sv.poll(64)}
    assert res[bad_id].status == dg.STATUS_CODES["DG_ERR_NUMERIC"]
    assert all(res[i].status == 0 for i in good_ids)


def c5_shaped_requests(n=10, seed=21, rate=40.0, kmin=1000, kmax=4000):
    """The serving configuration's recipe (Poisson arrivals, log-uniform mesh
    sizes, p in 1..3, one random vortex, end time uniform) at sizes and end
    times the oracle finishes in seconds."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        K = float(np.exp(rng.uniform(np.log(kmin), np.log(kmax))))
        nx = max(2, int(round(np.sqrt(K / 2.0))))
        p = int(rng.integers(1, 4))
        te = float(rng.uniform(0.05, 0.15))
        prob = dg.euler_vortex(((float(rng.uniform(2, 5)), float(rng.uniform(0, 10)), float(rng.uniform(0, 10))),))
        out.append(dg.request(nx, p, te, prob, mesh_seed=int(rng.integers(1, 2**31)), arrival_s=t))
    return out


def test_served_stream_vs_oracle():
    """Every served FP32 final state of a C5-shaped stream, with the larger
    requests on their own kernels inside the step graph and the others in the
    batch launch, is within 1e-5 relative L2 of the FP64 oracle run of the
    same request, and conserves mass to FP32 round-off."""
    import oracle as orc

    reqs = c5_shaped_requests()
    sizes = [2 * r.nx * r.ny for r in reqs]
    assert min(sizes) < 2000 <= max(sizes) and {r.p for r in reqs} == {1, 2, 3}
    sv = dg.Server(dg.DG_FP32, max_slots=6, steps_per_round=8, cfl=0.3, solo_min_elems=2000)
    sv.keep_states(True)
    t0 = sv.now()
    ids = []
    for r in reqs:
        r.arrival_s += t0
        ids.append(sv.submit(r))
    sv.drain()
    res = {r.id: r for r in sv.poll(64)}
    assert len(res) == len(reqs)
    for k, r in enumerate(reqs):
        out = res[ids[k]]
        assert out.status == 0 and out.t == r.t_end
        m = dg.Mesh.periodic_jittered(r.nx, r.ny, r.x0, r.x1, r.y0, r.y1, r.jitter, r.mesh_seed)
        o = orc.Oracle.from_mesh(m, r.p, r.prob)
        ref, n_ref, _ = o.integrate(o.project_ic(orc.ic_array(r.prob, m.period)), r.t_end, 0.3)
        got = sv.get_state(ids[k], (4, dg.n_modes(r.p), m.num_elem))
        err = orc.rel_l2(got, ref)
        print("request %d: %d triangles p=%d t_end %.3f, %d/%d steps, rel L2 vs oracle %.3e"
              % (k, m.num_elem, r.p, r.t_end, out.steps, n_ref, err), flush=True)
        assert err <= 1e-5
        m0, m1 = np.array(out.mass0), np.array(out.mass)
        assert np.all(np.abs(m1 - m0) <= 2e-6 * np.abs(m0) + 1e-9)
    sv.close()
