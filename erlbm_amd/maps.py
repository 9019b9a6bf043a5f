"""Batched tiny-lattice runs: the stability-map drivers on the GPU.

The reference's map drivers (src/main_stability-map_taylor-green.cpp,
src/main_stability-map_obstacle.cpp) loop over map points with
``omp parallel for`` and run ``roda_ponto`` on a private lattice per point.
Here a ``Batch`` holds the geometry once and ``run`` executes any number of
jobs (one per map point) in the persistent LDS-resident kernel of
``liberlbm.so`` (erlbm_batch_run in include/erlbm.h).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from ._lib import JOB_DTYPE, RESULT_DTYPE, check, lib


class Batch:
    """A shared D3Q19 tiny geometry (nz = 1, <= 256 fluid sites)."""

    def __init__(self, nx: int, ny: int, mask: Optional[np.ndarray] = None, inlet_x: int = -1,
                 outlet_x: int = -1, stream: int = 0):
        h = C.c_void_p()
        m = None if mask is None else np.ascontiguousarray(mask, np.int32).reshape(-1)
        mp = None if m is None else m.ctypes.data_as(C.c_void_p)
        check(lib().erlbm_batch_create(nx, ny, mp, inlet_x, outlet_x, C.c_void_p(stream or None), C.byref(h)),
              "erlbm_batch_create")
        self._h = h
        self.nx, self.ny = nx, ny

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().erlbm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_launch(self, wg_per_cu_f64: int = 0, wg_per_cu_f32: int = 0) -> None:
        check(lib().erlbm_batch_set_launch(self._h, wg_per_cu_f64, wg_per_cu_f32), "erlbm_batch_set_launch")

    def set_chunk(self, chunk_steps: int = 0) -> None:
        """0: one persistent launch per group to retirement; > 0: graph-replayed
        chunks of chunk_steps steps with early-exit compaction between them."""
        check(lib().erlbm_batch_set_chunk(self._h, chunk_steps), "erlbm_batch_set_chunk")

    def autotune(self, sample: np.ndarray, max_steps: int = 2000) -> dict:
        """Time the sample, every budget cut to max_steps, under every
        workgroups-per-CU value; keep and return the fastest."""
        sample = np.ascontiguousarray(sample, dtype=JOB_DTYPE)
        w64, w32 = C.c_int(), C.c_int()
        pps = C.c_double()
        check(lib().erlbm_batch_autotune(self._h, sample.ctypes.data_as(C.c_void_p), sample.size, max_steps,
                                         C.byref(w64), C.byref(w32), C.byref(pps)), "erlbm_batch_autotune")
        return {"wg_per_cu_f64": w64.value, "wg_per_cu_f32": w32.value, "sample_points_per_s": pps.value}

    def run(self, jobs: np.ndarray) -> np.ndarray:
        """jobs: structured array of JOB_DTYPE; returns RESULT_DTYPE results."""
        jobs = np.ascontiguousarray(jobs, dtype=JOB_DTYPE)
        res = np.zeros(jobs.size, dtype=RESULT_DTYPE)
        check(lib().erlbm_batch_run(self._h, jobs.ctypes.data_as(C.c_void_p), jobs.size,
                                    res.ctypes.data_as(C.c_void_p)), "erlbm_batch_run")
        return res

    def run_noalloc(self, jobs: np.ndarray) -> np.ndarray:
        """run() through erlbm_batch_run_noalloc: no device allocation, at
        most NOALLOC_MAX_JOBS jobs."""
        jobs = np.ascontiguousarray(jobs, dtype=JOB_DTYPE)
        res = np.zeros(jobs.size, dtype=RESULT_DTYPE)
        check(lib().erlbm_batch_run_noalloc(self._h, jobs.ctypes.data_as(C.c_void_p), jobs.size,
                                            res.ctypes.data_as(C.c_void_p)), "erlbm_batch_run_noalloc")
        return res


def tg_map_jobs(n_ma=400, n_nu=400, collision=L.COL_EXT_REG, tau_nh=0.65, precision=L.FP64, i0=0, i1=None, j0=0,
                j1=None) -> np.ndarray:
    i1 = n_ma if i1 is None else i1
    j1 = n_nu if j1 is None else j1
    jobs = np.zeros((j1 - j0) * (i1 - i0), dtype=JOB_DTYPE)
    check(lib().erlbm_tg_map_jobs(n_ma, n_nu, collision, tau_nh, precision, i0, i1, j0, j1,
                                  jobs.ctypes.data_as(C.c_void_p)), "erlbm_tg_map_jobs")
    return jobs


def obstacle_map_jobs(n_tau=400, n_nu=400, ma=0.1, precision=L.FP64, i0=0, i1=None, j0=0, j1=None) -> np.ndarray:
    i1 = n_tau if i1 is None else i1
    j1 = n_nu if j1 is None else j1
    jobs = np.zeros((j1 - j0) * (i1 - i0), dtype=JOB_DTYPE)
    check(lib().erlbm_obstacle_map_jobs(n_tau, n_nu, ma, precision, i0, i1, j0, j1,
                                        jobs.ctypes.data_as(C.c_void_p)), "erlbm_obstacle_map_jobs")
    return jobs


def tg_batch() -> Batch:
    """The Taylor-Green map lattice: 16 x 16 x 1, periodic."""
    return Batch(16, 16)


def obstacle_batch() -> Batch:
    """The obstacle map channel: 22 x 12 x 1, Zou-He inlet at x = 1, outlet at x = 20."""
    from .lattice import obstacle_mask
    return Batch(22, 12, obstacle_mask(), inlet_x=1, outlet_x=20)


def run_tg_map(collision=L.COL_EXT_REG, tau_nh=0.65, precision=L.FP64, n=400, batch: Optional[Batch] = None,
               **sub) -> np.ndarray:
    """Stability map (1 = unstable) as an (n_nu, n_ma) array, i fastest."""
    jobs = tg_map_jobs(n, n, collision, tau_nh, precision, **sub)
    own = batch is None
    b = tg_batch() if own else batch
    try:
        res = b.run(jobs)
    finally:
        if own:
            b.close()
    return res


def run_obstacle_map(precision=L.FP64, n=400, ma=0.1, batch: Optional[Batch] = None, **sub) -> np.ndarray:
    jobs = obstacle_map_jobs(n, n, ma, precision, **sub)
    own = batch is None
    b = obstacle_batch() if own else batch
    try:
        res = b.run(jobs)
    finally:
        if own:
            b.close()
    return res


def _boundary_distance(u: np.ndarray, r_max: int) -> np.ndarray:
    """Chessboard distance (capped at r_max + 1) of every map point to the
    nearest point whose 4-neighbourhood holds the other classification."""
    b = np.zeros(u.shape, bool)
    b[:-1, :] |= u[:-1, :] != u[1:, :]
    b[1:, :] |= u[1:, :] != u[:-1, :]
    b[:, :-1] |= u[:, :-1] != u[:, 1:]
    b[:, 1:] |= u[:, 1:] != u[:, :-1]
    dist = np.full(u.shape, r_max + 1, np.int32)
    reach = b.copy()
    for r in range(r_max + 1):
        dist[reach & (dist > r)] = r
        grown = reach.copy()   # one chessboard step further
        grown[1:, :] |= reach[:-1, :]
        grown[:-1, :] |= reach[1:, :]
        reach = grown.copy()
        reach[:, 1:] |= grown[:, :-1]
        reach[:, :-1] |= grown[:, 1:]
    return dist


def _margin_band(res, u64, s64, t64: float, dt: float, threshold: float) -> dict:
    """The other re-run rule: FP64 for the points whose result margin
    (max E/E0 over the energy tests, or max |rho| / dens_limit) lies closest
    to the stop threshold, |log(margin / threshold)| <= delta with delta the
    smallest that reaches 99.9% agreement (non-finite margins last).  Also
    where the disagreeing points' margins sit against all points that ran
    stable in the lower precision."""
    u = res["unstable"]
    m = res["margin"]
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.abs(np.log(m / threshold))
    c = np.where(np.isfinite(c), c, np.inf)
    order = np.argsort(c, kind="stable")
    wrong = (u != u64)[order]
    need = int(np.ceil(wrong.sum() - 0.001 * u.size))      # wrong points the band must fix
    fixed = np.cumsum(wrong)
    k = int(np.searchsorted(fixed, need) + 1) if need > 0 else 0
    delta = float(c[order[k - 1]]) if k else -1.0
    band = c <= delta                                        # every point as close as the last one needed
    k = int(band.sum())
    rerun = float(s64[band].astype(float).sum() / s64.astype(float).sum() * t64)
    stable = u == 0
    differ_stable = stable & (u != u64)
    q = lambda x: [round(float(v), 4) for v in np.quantile(x, [0.0, 0.5, 1.0])] if x.size else None
    return {"points": k, "delta": (round(delta, 4) if k else 0.0),
            "agreement": round(float(np.mean(np.where(band, u64, u) == u64)), 5),
            "rerun_seconds": round(rerun, 3), "total_seconds": round(dt + rerun, 3),
            "points_per_s": round(u.size / (dt + rerun), 1),
            "differing": int((u != u64).sum()), "differing_stable_here": int(differ_stable.sum()),
            "margin_quantiles_differing_stable": q(m[differ_stable]),
            "margin_quantiles_all_stable": q(m[stable])}


def precision_sweep(n: int = 400, r_max: int = 120) -> dict:
    """The three published maps in FP64, FP32-shifted storage (FP64
    arithmetic) and FP32 arithmetic on FP32-shifted storage: points/s (the
    batch autotuned per precision), classification agreement with the FP64
    map, and the cheapest mixed scheme that reaches 99.9% agreement -- the
    FP32 map, then the FP64 re-run of every point within r map cells of the
    FP32 map's stability boundary, r the smallest that reaches 99.9% with the
    FP64 map as the arbiter (a best case: a service would not know the FP64
    map); the re-run is priced at the FP64 map's time per node-step."""
    import time
    out = {}
    for name, mk, col, tnh in (("tg_erlbm065", tg_batch, L.COL_EXT_REG, 0.65), ("tg_reg", tg_batch, L.COL_REG, 1.0),
                               ("obstacle", obstacle_batch, L.COL_EXT_REG, 0.65)):
        runs = {}
        for pname, prec in (("fp64", L.FP64), ("fp32_storage", L.FP32_SHIFTED), ("fp32_arith", L.FP32_SHIFTED_F32)):
            jobs = (obstacle_map_jobs(n, n, 0.1, prec) if name == "obstacle" else tg_map_jobs(n, n, col, tnh, prec))
            with mk() as b:
                b.autotune(np.random.default_rng(3).choice(jobs, min(4096, jobs.size), replace=False), max_steps=2000)
                t = time.perf_counter()
                res = b.run(jobs)
                runs[pname] = (res, time.perf_counter() - t)
        r64, t64 = runs["fp64"]
        u64 = r64["unstable"].reshape(n, n)
        s64 = r64["steps"].astype(float).reshape(n, n)
        m = {"fp64": {"seconds": round(t64, 3), "points_per_s": round(n * n / t64, 1)}}
        for pname in ("fp32_storage", "fp32_arith"):
            res, dt = runs[pname]
            u = res["unstable"].reshape(n, n)
            e = {"seconds": round(dt, 3), "points_per_s": round(n * n / dt, 1),
                 "agreement_vs_fp64": round(float(np.mean(u == u64)), 5)}
            dist = _boundary_distance(u, r_max)
            for r in range(r_max + 1):
                band = dist <= r
                if np.mean(np.where(band, u64, u) == u64) >= 0.999:
                    rerun = float(s64[band].sum() / s64.sum() * t64)
                    e["fp64_rerun_band"] = {"radius": r, "points": int(band.sum()), "rerun_seconds": round(rerun, 3),
                                            "total_seconds": round(dt + rerun, 3),
                                            "points_per_s": round(n * n / (dt + rerun), 1)}
                    break
            e["fp64_rerun_margin"] = _margin_band(res, u64.reshape(-1), s64.reshape(-1), t64, dt,
                                                  1.0 if name == "obstacle" else 2.0)
            m[pname] = e
        out[name] = m
    return out


def main(argv=None) -> int:
    """python -m erlbm_amd.maps [--map tg|tg_reg|obstacle] [--n 400] [--precision fp64|fp32|auto]
    [--out FILE]: one full stability map, timed; prints one JSON line and
    optionally writes the map as the reference's map drivers do."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--map", default="tg", choices=["tg", "tg_reg", "obstacle"])
    ap.add_argument("--n", type=int, default=400)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "auto"])
    ap.add_argument("--out", default=None, help="write the map file (the reference drivers' VTK format)")
    ap.add_argument("--precision-sweep", action="store_true",
                    help="all three maps in FP64, FP32 storage and FP32 arithmetic: points/s and agreement")
    a = ap.parse_args(argv)
    if a.precision_sweep:
        print(json.dumps(precision_sweep(a.n)))
        return 0
    prec = {"fp64": L.FP64, "fp32": L.FP32_SHIFTED, "auto": L.FP_AUTO}[a.precision]
    if a.map == "obstacle":
        jobs, mk, nf = obstacle_map_jobs(a.n, a.n, 0.1, prec), obstacle_batch, 232
    else:
        col, tnh = (L.COL_REG, 1.0) if a.map == "tg_reg" else (L.COL_EXT_REG, 0.65)
        jobs, mk, nf = tg_map_jobs(a.n, a.n, col, tnh, prec), tg_batch, 256
    with mk() as b:
        t = time.perf_counter()
        res = b.run(jobs)
        dt = time.perf_counter() - t
    if a.out:
        from . import fields
        m = res["unstable"].reshape(a.n, a.n).astype(np.int32)   # (n_j, n_i), i fastest
        if a.map == "obstacle":
            fields.write_obstacle_map_vtk(a.out, m, 0.1)
        else:
            fields.write_tg_map_vtk(a.out, m, col, tnh)
    print(json.dumps({"map": a.map, "points": int(jobs.size), "seconds": round(dt, 3),
                      "points_per_s": round(jobs.size / dt, 1), "unstable": int(res["unstable"].sum()),
                      "node_steps_per_s": float(res["steps"].astype(float).sum() * nf / dt)}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
