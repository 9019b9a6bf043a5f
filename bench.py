#!/usr/bin/env python3
"""Benchmark: D2Q9 FP64 E-RLBM collide-stream, double shear layer L = 2048.

Workload (the configuration the headline metric is quoted on): D2Q9, periodic
L = 2048, E-RLBM tau_nh = 0.65, Re = 80000, U0 = 1/32, tau = 0.5 + 3 U0 L/Re,
rho = 1, equilibrium initialisation of the double-shear-layer field
(src/main_double-shear-layer.cpp:48-66, 87).  A step is one fused
pull-stream + collide sweep over all 4,194,304 nodes on one GPU, inputs
resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Prints ONE JSON line with the MLUPS, the roofline of the step kernel
(144 algorithmic bytes per node-step: 9 FP64 populations read + 9 written)
against MI355X's 8 TB/s, and the reference's own CPU step loop (compiled from
its sources into oracle/_ref, OpenMP on the host cores) timed on a bounded
sample as the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L = 2048
U0 = 1.0 / 32.0
RE = 80000.0
TAU = 0.5 + 3.0 * U0 * L / RE
TAU_NH = 0.65
BYTES_PER_NODE = 144           # 9 x 8 B read + 9 x 8 B written
HBM_PEAK_GBS = 8000.0
# rocprofv3 PMC traffic per launch (FETCH_SIZE + WRITE_SIZE, calibrated against
# a plain copy kernel); filled from profiles/ when a measurement exists.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "dsl2048_d2q9_fp64_traffic.json")


def cpu_baseline(seconds: float):
    """The reference's OpenMP step loop on a bounded sample of the same
    workload (the first n steps of the L=2048 run), all host cores."""
    import oracle
    n_threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    f0 = oracle.Oracle().init_double_shear_d2q9(L)
    if oracle.Reference.available(9, fma=True):
        ref = oracle.Reference(9, fma=True)
        kind = "reference"
        run = lambda f, n: ref.run_steps_omp(L, L, 1, f, n, oracle.COL_EXT_REG, TAU, TAU_NH, n_threads)  # noqa: E731
    else:
        o = oracle.Oracle(fma=True)
        kind = "port"
        ini, st, nf = o.stream_map(9, L, L, 1)

        def run(f, n):
            t = time.perf_counter()
            o.run_steps(9, f, st, n, oracle.COL_EXT_REG, TAU, TAU_NH, nthreads=n_threads)
            return time.perf_counter() - t
    t1 = run(f0.copy(), 1)
    n = max(2, min(2000, int(seconds / max(t1, 1e-6))))
    dt = run(f0.copy(), n)
    return {"value": L * L * n / dt / 1e6, "unit": "MLUPS", "cores": n_threads, "kind": kind,
            "sample": f"D2Q9 E-RLBM double shear L={L}, first {n} steps ({dt:.1f} s wall, "
                      f"{'reference step loop built from its sources, -O3 -march=x86-64-v3, OpenMP' if kind == 'reference' else 'C restatement, OpenMP'})"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "fp16"],
                    help="population storage of the timed run (fp32/fp16: shifted deviations)")
    ap.add_argument("--copy-only", action="store_true",
                    help="run only the calibration copy kernel (302 MB read + 302 MB written per launch)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        return 2

    import erlbm_amd as E

    if E.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    # plain copy of the bytes one step moves (half read, half written)
    copy_bytes = BYTES_PER_NODE * L * L // 2
    copy_ms = E.hbm_copy_bench(copy_bytes, 5 if a.copy_only else 20)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    if a.copy_only:
        print(json.dumps({"copy_bytes_each_way": copy_bytes, "ms": copy_ms, "GBs": copy_gbs}))
        return 0
    prec = {"fp64": E.FP64, "fp32": E.FP32_SHIFTED, "fp16": E.FP16_SHIFTED}[a.precision]
    bytes_per_node = {"fp64": 144, "fp32": 72, "fp16": 36}[a.precision]
    f0 = E.init_double_shear(L)
    lat = E.Lattice(9, L, L, 1, precision=prec)
    lat.set_params(TAU, TAU_NH, 1.0, E.COL_EXT_REG)
    lat.upload(f0)
    mass0 = float(f0.sum())
    del f0
    lat.step(a.warmup)
    lat.synchronize()
    ms = lat.step_timed(a.steps)          # HIP events on the lattice's stream
    steps_done, reason = lat.status()
    ms_per_step = ms / a.steps
    mlups = L * L / (ms_per_step * 1e-3) / 1e6
    achieved = bytes_per_node * L * L / (ms_per_step * 1e-3) / 1e9
    post = lat.download(E.POST_COLLISION)
    mass_err = abs(float(post.sum()) - mass0) / mass0
    finite = bool(np.isfinite(post).all())
    del post
    lat.close()
    if steps_done != a.warmup + a.steps or reason != 0 or not finite or mass_err > (1e-9 if prec == E.FP64 else 1e-3):
        print(f"bench.py: run check failed (steps {steps_done}, stop {reason}, finite {finite}, "
              f"mass error {mass_err:.2e})", file=sys.stderr)
        return 1

    traffic = None
    if a.precision == "fp64" and os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_seconds)
    line = {
        "metric": "D2Q9 FP64 MLUPS (%HBM peak); map points/s + p99 latency, FP32 vs FP64",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (shifted storage, f32 arithmetic)",
                  "fp16": "f16 storage / f32 arithmetic (shifted)"}[a.precision],
        "data": "synthetic: double-shear-layer initial field (deterministic, no dataset)",
        "config": {"workload": "D2Q9 E-RLBM double shear layer, periodic L=2048 (4,194,304 nodes), "
                               f"tau=0.5024, tau_nh=0.65, {a.precision} row-blocked pull-stream+collide",
                   "L": L, "tau": TAU, "tau_nh": TAU_NH, "collision": "E-RLBM", "precision": a.precision},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "copy_kernel_GBs": round(copy_gbs, 1),
        "mass_rel_error": mass_err,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
