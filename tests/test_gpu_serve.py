"""Serving loop: continuous batching of uneven solves.

Every response must equal the same solve run alone (bitwise for FP64 map
points and D2Q9 runs: per-request arithmetic does not depend on what shares
the device), requests arriving while others run are admitted as memory frees
up, and the HBM budget is never exceeded.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _mix(S, n=40, seed=7, rate=400.0):
    return S.poisson_mix(n, rate, seed=seed, dsl_t_final=0.02, tg_sizes=(64, 128))


def _alone(gpu, S, solve):
    """The same request through the batch API (map points) or the plain
    lattice path (D2Q9 runs), alone."""
    r = solve.req
    if r.kind == S.REQ_TG_MAP_POINT:
        jobs = gpu.tg_map_jobs(r.n_i, r.n_j, r.collision, r.tau_nh, r.precision, r.i, r.i + 1, r.j, r.j + 1)
        with gpu.tg_batch() as b:
            res = b.run(jobs)[0]
        return int(res["unstable"]), int(res["steps"]), [float(res["margin"]), float(res["e0"]), float(res["energy"])]
    if r.kind == S.REQ_OBSTACLE_MAP_POINT:
        jobs = gpu.obstacle_map_jobs(r.n_i, r.n_j, 0.1, r.precision, r.i, r.i + 1, r.j, r.j + 1)
        with gpu.obstacle_batch() as b:
            res = b.run(jobs)[0]
        return int(res["unstable"]), int(res["steps"]), [float(res["margin"]), float(res["e0"]), float(res["energy"])]
    one = S.run_plain([S.Solve(req=r, t_arrival=0.0)])
    return None, None, one


def test_served_results_equal_alone(gpu):
    from erlbm_amd import serve as S
    solves = _mix(S)
    summary = S.run_served(solves, chunk_steps=500)
    assert summary["all_ok"]
    kinds = {s.response["kind"] for s in solves}
    print(f"served {summary['requests']} requests of kinds {sorted(kinds)}: "
          f"p50 {summary['p50_s']:.3f}s p99 {summary['p99_s']:.3f}s", flush=True)
    for s in solves:
        got = s.response
        if s.req.kind in (S.REQ_TG_MAP_POINT, S.REQ_OBSTACLE_MAP_POINT):
            u, st, val = _alone(gpu, S, s)
            assert (got["unstable"], got["steps"]) == (u, st), (s.req.kind, s.req.i, s.req.j)
            assert got["value"][0] == val[0] or (np.isnan(got["value"][0]) and np.isnan(val[0]))
        else:
            S.run_plain([s2 := S.Solve(req=s.req, t_arrival=0.0)])
            want = s2.response
            assert got["steps"] == want["steps"]
            assert got["value"] == want["value"], (got, want)


def test_chunking_does_not_change_results(gpu):
    from erlbm_amd import serve as S
    a = _mix(S, n=24, seed=11)
    b = _mix(S, n=24, seed=11)
    S.run_served(a, chunk_steps=100)
    S.run_served(b, chunk_steps=2000)
    for x, y in zip(a, b):
        assert (x.response["unstable"], x.response["steps"], x.response["value"]) == \
               (y.response["unstable"], y.response["steps"], y.response["value"])


def test_admission_waits_for_budget(gpu):
    """A budget that holds one pool and little else: D2Q9 runs queue behind
    each other, everything still completes, memory stays within budget."""
    from erlbm_amd import serve as S
    solves = _mix(S, n=16, seed=3)
    # one pool of 64 FP64 Taylor-Green slots (~2.6 MB) + one pool of obstacle slots + one L=128 lattice
    summary = S.run_served(solves, budget_bytes=12 << 20, chunk_steps=500, tiny_slots=64)
    assert summary["all_ok"], summary
    t_wait = [s.response["t_admit"] - s.response["t_submit"] for s in solves]
    print(f"max admission wait {max(t_wait):.3f}s", flush=True)


def _stated_mix(S, gpu, n=32, seed=20261203):
    """The serving mix at its stated sizes (Taylor-Green L in {128, 256,
    512}, double shear L = 512 to t* = 0.25), seeded, with one request of
    each D2Q9 size written over early arrivals so every kind and size is in
    the mix whatever the draw."""
    solves = S.poisson_mix(n, 40.0, seed=seed)
    for k, (kind, Ls) in zip((1, 4, 7, 10), ((S.REQ_TAYLOR_GREEN, 128), (S.REQ_TAYLOR_GREEN, 256),
                                              (S.REQ_TAYLOR_GREEN, 512), (S.REQ_DOUBLE_SHEAR, 512))):
        r = solves[k].req
        r.kind, r.L, r.precision, r.collision, r.tau_nh = kind, Ls, gpu.FP64, gpu.COL_EXT_REG, 0.65
        r.t_final = 0.25
    kinds = {s.req.kind for s in solves}
    assert kinds == {S.REQ_TG_MAP_POINT, S.REQ_OBSTACLE_MAP_POINT, S.REQ_TAYLOR_GREEN, S.REQ_DOUBLE_SHEAR}, kinds
    return solves


def test_stated_mix_served_equals_alone_and_oracle(gpu, orc):
    """Continuous batching of the stated mix: every response equals the same
    request run alone (bitwise), and each Taylor-Green L = 128 response's
    E_vector, E_magnitude and (nu_eff - nu)/nu equal those of the oracle's
    own run of src/main_taylor-green.cpp:104-217."""
    from erlbm_amd import serve as S
    solves = _stated_mix(S, gpu)
    summary = S.run_served(solves, chunk_steps=1000)
    assert summary["all_ok"], summary
    print(f"stated mix: {summary['requests']} requests, p50 {summary['p50_s']:.3f}s p99 {summary['p99_s']:.3f}s, "
          f"kinds {summary['per_kind']}", flush=True)
    tg128 = None
    for s in solves:
        got = s.response
        print(f"  {got['kind']} L={s.req.L} i={s.req.i} j={s.req.j}: steps {got['steps']}", flush=True)
        if s.req.kind in (S.REQ_TG_MAP_POINT, S.REQ_OBSTACLE_MAP_POINT):
            u, st, val = _alone(gpu, S, s)
            assert (got["unstable"], got["steps"]) == (u, st), (s.req.kind, s.req.i, s.req.j)
            assert got["value"][0] == val[0] or (np.isnan(got["value"][0]) and np.isnan(val[0]))
        else:
            S.run_plain([s2 := S.Solve(req=s.req, t_arrival=0.0)])
            want = s2.response
            assert got["steps"] == want["steps"] and got["unstable"] == want["unstable"] == 0
            assert got["value"] == want["value"], (got, want)
            if s.req.kind == S.REQ_TAYLOR_GREEN and s.req.L == 128:
                tg128 = got
    assert tg128 is not None
    ref = orc.run_taylor_green_d2q9(L=128, col=O.COL_EXT_REG, tau_nh=0.65, nthreads=16)
    print(f"TG L=128 served E_vector {tg128['value'][0]:.9g}, oracle {ref[0]:.9g}", flush=True)
    assert tg128["steps"] == int(ref[4]) + 1
    assert list(tg128["value"]) == list(ref[:3])


def test_request_larger_than_budget_is_answered(gpu):
    """A request that needs more than the whole HBM budget (a double shear
    L = 512 lattice, 44 MB, under 12 MiB) is answered ERLBM_ERR_NOMEM at
    once instead of waiting for ever; the map points around it are served."""
    from erlbm_amd import serve as S
    from erlbm_amd._lib import ERLBM_ERR_NOMEM
    solves = S.poisson_mix(8, 400.0, seed=5, tg_sizes=(64,))
    for s in solves:
        s.req.kind = S.REQ_TG_MAP_POINT
    r = solves[3].req
    r.kind, r.L, r.precision, r.t_final = S.REQ_DOUBLE_SHEAR, 512, gpu.FP64, 0.02
    summary = S.run_served(solves, budget_bytes=12 << 20, chunk_steps=500, tiny_slots=64)
    assert solves[3].response["status"] == ERLBM_ERR_NOMEM
    assert all(s.response["status"] == 0 for k, s in enumerate(solves) if k != 3)
    assert not summary["all_ok"]


def test_idle_pool_released_for_a_waiting_request(gpu):
    """Under a budget that holds one pool of map slots and little else, an
    obstacle map point that arrives after the Taylor-Green points have
    retired is admitted once the idle Taylor-Green pool is freed."""
    from erlbm_amd import serve as S
    first = S.poisson_mix(4, 400.0, seed=9, tg_sizes=(64,))
    for s in first:
        s.req.kind = S.REQ_TG_MAP_POINT
    later = S.poisson_mix(2, 400.0, seed=10, tg_sizes=(64,))
    for s in later:
        s.req.kind = S.REQ_OBSTACLE_MAP_POINT
    # a pool of 64 slots holds 2.5 MB (Taylor-Green) / 2.6 MB (obstacle)
    with S.Server(budget_bytes=4 << 20, chunk_steps=500, tiny_slots=64) as srv:
        used_one = None
        for batch in (first, later):
            ids = {srv.submit(s.req): s for s in batch}
            while any(s.response is None for s in batch):
                srv.pump()
                for r in srv.poll():
                    ids[r["id"]].response = r
            if used_one is None:
                used_one = srv.memory()[0]
        used, budget = srv.memory()
    assert all(s.response["status"] == 0 for s in first + later)
    assert used <= budget and 2 * used_one > budget    # the two pools never fitted together
