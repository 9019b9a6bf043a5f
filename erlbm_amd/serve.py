"""Serving: the continuous-batching request loop (erlbm_server_* in
include/erlbm.h) and the one-run-at-a-time plain FP64 path it is measured
against, over a seeded Poisson stream of the four reference drivers' solves.

    python -m erlbm_amd.serve [--n 200] [--rate 40] [--seed 20261203]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib as L
from ._lib import check, lib

REQ_TG_MAP_POINT, REQ_OBSTACLE_MAP_POINT, REQ_TAYLOR_GREEN, REQ_DOUBLE_SHEAR = 1, 2, 3, 4
KIND_NAMES = {1: "tg_map_point", 2: "obstacle_map_point", 3: "taylor_green", 4: "double_shear"}


from ._lib import Request, Response  # noqa: E402


def _bind():
    return lib()


@dataclass
class Solve:
    """One request of the stream and, once served, its outcome."""
    req: Request
    t_arrival: float
    id: int = -1
    response: Optional[dict] = None


def resp_dict(r: Response) -> dict:
    return {"id": r.id, "kind": KIND_NAMES.get(r.kind, r.kind), "status": r.status, "unstable": r.unstable,
            "precision": r.precision, "steps": r.steps, "value": list(r.value), "t_submit": r.t_submit,
            "t_admit": r.t_admit, "t_done": r.t_done}


class Server:
    def __init__(self, budget_bytes: int = 32 << 30, chunk_steps: int = 1000, tiny_slots: int = 4096):
        self.lb = _bind()
        h = C.c_void_p()
        check(self.lb.erlbm_server_create(budget_bytes, chunk_steps, tiny_slots, None, C.byref(h)),
              "erlbm_server_create")
        self._h = h

    def close(self):
        if self._h.value:
            self.lb.erlbm_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req: Request) -> int:
        i = C.c_longlong()
        check(self.lb.erlbm_server_submit(self._h, C.byref(req), C.byref(i)), "erlbm_server_submit")
        return i.value

    def pump(self):
        a, b = C.c_int(), C.c_int()
        check(self.lb.erlbm_server_pump(self._h, C.byref(a), C.byref(b)), "erlbm_server_pump")
        return a.value, b.value

    def poll(self, max_out: int = 4096) -> List[dict]:
        buf = (Response * max_out)()
        n = C.c_int()
        check(self.lb.erlbm_server_poll(self._h, buf, max_out, C.byref(n)), "erlbm_server_poll")
        return [resp_dict(buf[k]) for k in range(n.value)]

    def clock(self) -> float:
        t = C.c_double()
        check(self.lb.erlbm_server_clock(self._h, C.byref(t)), "erlbm_server_clock")
        return t.value

    def memory(self):
        """(bytes of budget in use, budget)."""
        used, budget = C.c_longlong(), C.c_longlong()
        check(self.lb.erlbm_server_memory(self._h, C.byref(used), C.byref(budget)), "erlbm_server_memory")
        return used.value, budget.value


def poisson_mix(n: int, rate: float, seed: int = 20261203, map_precision: int = L.FP64, dsl_t_final: float = 0.25,
                tg_sizes=(128, 256, 512)) -> List[Solve]:
    """Seeded Poisson arrivals: 60% Taylor-Green map points, 25% obstacle map
    points, 10% D2Q9 Taylor-Green at L in tg_sizes, 5% D2Q9 double shear at
    L = 512 (t* = dsl_t_final).  Map points use map_precision; the D2Q9
    accuracy runs use FP64."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        u = rng.uniform()
        r = Request()
        r.n_i = r.n_j = 400
        r.tau_nh = 0.65
        r.collision = L.COL_EXT_REG
        if u < 0.60:
            r.kind = REQ_TG_MAP_POINT
            r.i, r.j = int(rng.integers(400)), int(rng.integers(400))
            r.precision = map_precision
        elif u < 0.85:
            r.kind = REQ_OBSTACLE_MAP_POINT
            r.i, r.j = int(rng.integers(400)), int(rng.integers(400))
            r.precision = map_precision
        elif u < 0.95:
            r.kind = REQ_TAYLOR_GREEN
            r.L = int(rng.choice(tg_sizes))
            r.precision = L.FP64
        else:
            r.kind = REQ_DOUBLE_SHEAR
            r.L = 512
            r.precision = L.FP64
            r.t_final = dsl_t_final
        out.append(Solve(req=r, t_arrival=t))
    return out


def run_served(solves: List[Solve], budget_bytes: int = 32 << 30, chunk_steps: int = 1000,
               tiny_slots: int = 4096, time_scale: float = 1.0) -> dict:
    """Replay the arrival stream against the server (continuous batching)."""
    with Server(budget_bytes, chunk_steps, tiny_slots) as srv:
        t0 = srv.clock()
        pending = list(range(len(solves)))
        by_id = {}
        done = 0
        while done < len(solves):
            now = srv.clock() - t0
            while pending and solves[pending[0]].t_arrival * time_scale <= now:
                k = pending.pop(0)
                solves[k].id = srv.submit(solves[k].req)
                by_id[solves[k].id] = k
            in_flight, queued = srv.pump()
            for r in srv.poll():
                solves[by_id[r["id"]]].response = r
                done += 1
            if in_flight == 0 and queued == 0 and pending:
                # idle: wait for the next arrival
                wait = solves[pending[0]].t_arrival * time_scale - (srv.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
        t_end = srv.clock() - t0
    return summarize(solves, t_end, "served")


def run_plain(solves: List[Solve], time_scale: float = 1.0) -> dict:
    """The same stream served one run at a time through the plain FP64
    lattice path (one lattice, one kernel launch per step)."""
    from .lattice import (Lattice, init_double_shear, init_taylor_green, init_tg_map, obstacle_map_point_params,
                          obstacle_mask, taylor_green_error, tg_map_point_params)
    t0 = time.perf_counter()
    for s in solves:
        now = time.perf_counter() - t0
        if s.t_arrival * time_scale > now:
            time.sleep(s.t_arrival * time_scale - now)
        t_start = time.perf_counter() - t0
        r = s.req
        val = [0.0, 0.0, 0.0]
        if r.kind == REQ_TG_MAP_POINT:
            u0, tau, tnh, n = tg_map_point_params(r.i, r.j, r.n_i, r.n_j, 16, r.collision, r.tau_nh)
            with Lattice(19, 16, 16, 1) as lat:
                lat.set_params(tau, tnh, 1.0, r.collision)
                lat.upload(init_tg_map(16, u0))
                lat.step(n, energy_interval=100, energy_factor=2.0)
                steps, reason = lat.status()
        elif r.kind == REQ_OBSTACLE_MAP_POINT:
            u, tau, tnh, n = obstacle_map_point_params(r.i, r.j, r.n_i, r.n_j)
            with Lattice(19, 22, 12, 1, mask=obstacle_mask()) as lat:
                lat.set_params(tau, tnh, 1.0, L.COL_EXT_REG)
                lat.set_zouhe_inlet(1, u)
                lat.set_outlet_rescale(20, 1.0)
                lat.set_density_check(10.0)
                from .lattice import dist_eq
                lat.upload(np.tile(dist_eq(19, 0.0, 0.0, 0.0, 1.0), (lat.n_fluid, 1)))
                lat.step(n)
                steps, reason = lat.status()
        elif r.kind == REQ_TAYLOR_GREEN:
            Ls = r.L
            u0 = 2.0 / Ls
            visc = u0 * Ls / 50.0
            tau = 0.5 + 3.0 * visc
            nT = int(round(Ls / u0))
            with Lattice(9, Ls, Ls, 1) as lat:
                lat.set_params(tau, 0.65, 1.0, r.collision)
                lat.upload(init_taylor_green(9, Ls, u0, tau))
                lat.step(nT + 1)
                steps, reason = lat.status()
                val = list(taylor_green_error(lat.download(L.POST_COLLISION), Ls, u0, visc, nT))
        else:
            Ls = r.L
            tau = 0.5 + 3.0 * (1.0 / 32.0) * Ls / 80000.0
            n = int(round(r.t_final * round(Ls / (1.0 / 32.0)))) + 1
            with Lattice(9, Ls, Ls, 1) as lat:
                lat.set_params(tau, 0.65, 1.0, r.collision)
                lat.upload(init_double_shear(Ls))
                lat.step(n)
                steps, reason = lat.status()
                val = [lat.kinetic_energy(L.POST_COLLISION, ordered=False), 0.0, 0.0]
        t_done = time.perf_counter() - t0
        s.response = {"kind": KIND_NAMES[r.kind], "unstable": int(reason != 0), "steps": steps, "value": val,
                      "t_submit": s.t_arrival * time_scale, "t_admit": t_start, "t_done": t_done, "status": 0,
                      "precision": L.FP64}
    return summarize(solves, time.perf_counter() - t0, "plain")


def summarize(solves: List[Solve], t_end: float, label: str) -> dict:
    lat = np.array([s.response["t_done"] - s.response["t_submit"] for s in solves])
    ok = all(s.response["status"] == 0 for s in solves)
    per_kind = {}
    for k in KIND_NAMES.values():
        sel = np.array([s.response["kind"] == k for s in solves])
        if sel.any():
            per_kind[k] = {"n": int(sel.sum()), "p50_s": float(np.percentile(lat[sel], 50)),
                           "p99_s": float(np.percentile(lat[sel], 99))}
    return {"mode": label, "requests": len(solves), "all_ok": ok, "wall_s": t_end,
            "throughput_req_per_s": len(solves) / t_end, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "per_kind": per_kind}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--rate", type=float, default=40.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--chunk", type=int, default=1000)
    ap.add_argument("--map-precision", default="fp64", choices=["fp64", "fp32", "auto"])
    ap.add_argument("--no-plain", action="store_true")
    a = ap.parse_args()
    mp = {"fp64": L.FP64, "fp32": L.FP32_SHIFTED, "auto": L.FP_AUTO}[a.map_precision]
    served = run_served(poisson_mix(a.n, a.rate, a.seed, mp), chunk_steps=a.chunk)
    out = {"served": served}
    if not a.no_plain:
        out["plain"] = run_plain(poisson_mix(a.n, a.rate, a.seed, L.FP64))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
