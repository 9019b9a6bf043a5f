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
