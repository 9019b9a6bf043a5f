"""Batched stability-map path against the reference's own map runs.

FP64 jobs are bitwise faithful, so their classification must equal the
strict reference build point for point (tests/golden/ref_map_subsample.npz:
roda_ponto of both map drivers on the README's 20 x 20 subsample positions).
FP32-shifted jobs are an approximation: their agreement is measured and
bounded below.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _subsample_jobs(gpu, kind, g, precision, col=O.COL_EXT_REG, tau_nh=0.65):
    rows = []
    for i, j in zip(g["i"], g["j"]):
        if kind == "tg":
            rows.append(gpu.tg_map_jobs(400, 400, col, tau_nh, precision, int(i), int(i) + 1, int(j), int(j) + 1))
        else:
            rows.append(gpu.obstacle_map_jobs(400, 400, 0.1, precision, int(i), int(i) + 1, int(j), int(j) + 1))
    return np.concatenate(rows)


@pytest.mark.parametrize("key,col,tnh", [("tg_erlbm065", O.COL_EXT_REG, 0.65), ("tg_reg", O.COL_REG, 1.0)])
def test_tg_map_subsample_fp64_equals_reference(gpu, golden, key, col, tnh):
    g = golden("ref_map_subsample.npz")
    jobs = _subsample_jobs(gpu, "tg", g, gpu.FP64, col, tnh)
    print(f"tg map subsample {key}: {jobs.size} points", flush=True)
    with gpu.tg_batch() as b:
        res = b.run(jobs)
    assert np.all(res["done"] == 1)
    agree = int(np.sum(res["unstable"] == g[f"{key}_strict"]))
    assert agree == jobs.size, f"{agree}/{jobs.size}"


def test_obstacle_map_subsample_fp64_equals_reference(gpu, golden):
    g = golden("ref_map_subsample.npz")
    jobs = _subsample_jobs(gpu, "obstacle", g, gpu.FP64)
    print(f"obstacle map subsample: {jobs.size} points", flush=True)
    with gpu.obstacle_batch() as b:
        res = b.run(jobs)
    agree = int(np.sum(res["unstable"] == g["obstacle_strict"]))
    assert agree == jobs.size, f"{agree}/{jobs.size}"


@pytest.mark.parametrize("key", ["tg_erlbm065", "tg_reg", "obstacle"])
def test_points_off_the_published_maps_equal_reference(gpu, golden, key):
    """The points where the full FP64 maps differ from the published data
    (185, 190 and 52 of 160,000; tests/golden/ref_map_disagreements.npz):
    there the GPU classifies every point as the reference's own strict build
    does (its FMA build splits about evenly: stability-edge points)."""
    g = golden("ref_map_disagreements.npz")
    idx = g[f"{key}_idx"]
    kind = "obstacle" if key == "obstacle" else "tg"
    col, tnh = (O.COL_REG, 1.0) if key == "tg_reg" else (O.COL_EXT_REG, 0.65)
    pts = {"i": idx % 400, "j": idx // 400}
    jobs = _subsample_jobs(gpu, kind, pts, gpu.FP64, col, tnh)
    print(f"{key}: {jobs.size} points off the published map", flush=True)
    with (gpu.obstacle_batch() if kind == "obstacle" else gpu.tg_batch()) as b:
        res = b.run(jobs)
    assert np.array_equal(res["unstable"].astype(np.uint8), g[f"{key}_strict"])
    assert np.array_equal(res["unstable"].astype(np.uint8), g[f"{key}_gpu"])


def test_stop_steps_and_energies_match_oracle(gpu, orc):
    """Stop step and energies of single points, bitwise, against the oracle's
    restatement of roda_ponto."""
    pts = [(380, 5), (150, 30), (20, 390), (399, 0)]
    jobs = np.concatenate([gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP64, i, i + 1, j, j + 1)
                           for i, j in pts])
    # cap the long stable runs to keep the CPU side short
    jobs["n_steps"] = np.minimum(jobs["n_steps"], 3000)
    with gpu.tg_batch() as b:
        res = b.run(jobs)
    for k, (i, j) in enumerate(pts):
        u0, tau, tnh = orc.tg_map_params(i, j)
        f0 = orc.init_tg_map_d3q19(16, u0)
        ini, st, nf = orc.stream_map(19, 16, 16, 1)
        e0 = orc.kinetic_energy(19, f0)
        assert res["e0"][k] == e0
        # replay the driver loop on the oracle up to the GPU's stop step
        a = f0
        stop = int(jobs["n_steps"][k])
        unstable = 0
        e_last = e0
        for s in range(int(jobs["n_steps"][k])):
            if s % 100 == 0 and s > 0:
                e_last = orc.kinetic_energy(19, a)
                if e_last > 2 * e0 or not np.isfinite(e_last):
                    stop, unstable = s, 1
                    break
            a = orc.run_steps(19, a, st, 1, O.COL_EXT_REG, tau, tnh)
        assert (res["unstable"][k], res["steps"][k]) == (unstable, stop), (i, j)
        if unstable:
            assert res["energy"][k] == e_last or (np.isnan(res["energy"][k]) and np.isnan(e_last))


def test_obstacle_points_match_oracle_steps(gpu, orc):
    pts = [(0, 0), (10, 20), (200, 3), (5, 399)]
    jobs = np.concatenate([gpu.obstacle_map_jobs(400, 400, 0.1, gpu.FP64, i, i + 1, j, j + 1) for i, j in pts])
    jobs["n_steps"] = np.minimum(jobs["n_steps"], 4000)
    with gpu.obstacle_batch() as b:
        res = b.run(jobs)
    m = orc.obstacle_mask()
    ini, st, nf = orc.stream_map(19, 22, 12, 1, m)
    for k, (i, j) in enumerate(pts):
        u, tau, tnh = orc.obstacle_map_params(i, j)
        out, n, v = orc.run_bc_steps(19, orc.init_rest_d3q19(nf), 22, 12, 1, ini, st, int(jobs["n_steps"][k]),
                                     O.COL_EXT_REG, tau, tnh, inlet_x=1, u_in=u, outlet_x=20, rho_out=1.0,
                                     dens_limit=10.0)
        assert (res["unstable"][k], res["steps"][k]) == (v, n), (i, j)


@pytest.mark.parametrize("prec_name,floor", [("FP_AUTO", 0.95), ("FP32_SHIFTED_F32", 0.85)])
def test_fp32_shifted_agreement(gpu, golden, prec_name, floor):
    """FP32-shifted storage (FP_AUTO: FP64 arithmetic, FP64 storage below
    tau - 1/2 < 1e-3; FP32_SHIFTED_F32: FP32 arithmetic everywhere) against
    the strict reference classification on the subsample: a regression floor
    for the non-default formats, not a parity bar (measured 1.000 / 0.998 /
    0.963 for FP_AUTO; 0.92 / 0.96 / 0.955 in FP32 arithmetic).  The
    reference's own strict and FMA builds disagree on 0.5-1.25% of these
    points: stability-edge points whose side is set by round-off."""
    g = golden("ref_map_subsample.npz")
    prec = getattr(gpu, prec_name)
    for key, kind, col, tnh in (("tg_erlbm065", "tg", O.COL_EXT_REG, 0.65), ("tg_reg", "tg", O.COL_REG, 1.0),
                                ("obstacle", "obstacle", O.COL_EXT_REG, 0.65)):
        jobs = _subsample_jobs(gpu, kind, g, prec, col, tnh)
        with (gpu.tg_batch() if kind == "tg" else gpu.obstacle_batch()) as b:
            res = b.run(jobs)
        agree = float(np.mean(res["unstable"] == g[f"{key}_strict"]))
        n32 = int(np.sum(res["precision"] != gpu.FP64))
        print(f"{key}: {prec_name} agreement {agree:.4f} ({n32} points in FP32 storage)", flush=True)
        assert n32 > 0
        assert agree >= floor, (key, agree)


def test_result_independent_of_batch_mix(gpu):
    """A job's result is the same alone and inside any batch (per-lattice
    arithmetic and reduction order do not depend on the batch)."""
    mix = gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP_AUTO, 100, 140, 150, 160)
    mix["n_steps"] = np.minimum(mix["n_steps"], 5000)
    with gpu.tg_batch() as b:
        together = b.run(mix)
        for k in (0, 7, 123, 399):
            alone = b.run(mix[k:k + 1])
            assert alone.tobytes() == together[k:k + 1].tobytes()
        rev = b.run(mix[::-1].copy())[::-1]
        assert rev.tobytes() == together.tobytes()


def _mixed_jobs(gpu):
    """Stable and unstable Taylor-Green points (RLBM and E-RLBM, FP64 and FP32
    storage) with budgets cut so the test stays short."""
    a = gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP64, 0, 400, 10, 11)
    b = gpu.tg_map_jobs(400, 400, O.COL_REG, 1.0, gpu.FP_AUTO, 0, 400, 300, 301)
    jobs = np.concatenate([a, b])
    jobs["n_steps"] = np.minimum(jobs["n_steps"], 3000)
    return jobs


def test_chunked_graph_mode_equals_one_launch(gpu):
    """Graph-replayed 100-step chunks with compaction of the live jobs give
    exactly the results of one persistent launch per group."""
    jobs = _mixed_jobs(gpu)
    obst = gpu.obstacle_map_jobs(400, 400, 0.1, gpu.FP64, 0, 400, 200, 201)
    obst["n_steps"] = np.minimum(obst["n_steps"], 4000)
    with gpu.tg_batch() as b:
        ref = b.run(jobs)
        for chunk in (100, 1000, 7):
            b.set_chunk(chunk)
            got = b.run(jobs)
            assert got.tobytes() == ref.tobytes(), chunk
    with gpu.obstacle_batch() as b:
        ref = b.run(obst)
        b.set_chunk(100)
        assert b.run(obst).tobytes() == ref.tobytes()
    assert 0 < int(ref["unstable"].sum()) < len(ref)


def test_batch_autotune_keeps_results(gpu):
    jobs = _mixed_jobs(gpu)
    with gpu.tg_batch() as b:
        ref = b.run(jobs)
        pick = b.autotune(jobs, max_steps=500)
        print(f"batch autotune: {pick}", flush=True)
        assert pick["wg_per_cu_f64"] >= 1 and pick["wg_per_cu_f32"] >= 1
        assert pick["sample_points_per_s"] > 0
        assert b.run(jobs).tobytes() == ref.tobytes()


def test_batch_edge_cases(gpu):
    """An empty job list, zero-step jobs and a single job; invalid jobs are
    rejected before anything runs."""
    with gpu.tg_batch() as b:
        assert b.run(gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP64, 0, 0, 0, 0)).size == 0
        jobs = gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP64, 10, 13, 20, 21)
        jobs["n_steps"][:] = 0
        res = b.run(jobs)
        assert (res["steps"] == 0).all() and (res["unstable"] == 0).all() and (res["done"] == 1).all()
        one = gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, gpu.FP64, 399, 400, 0, 1)
        one["n_steps"] = 300
        assert b.run(one)["steps"][0] in (100, 200, 300)
        bad = one.copy()
        bad["tau"] = -1.0
        with pytest.raises(RuntimeError, match="invalid job"):
            b.run(bad)
        bad = one.copy()
        bad["collision"] = O.COL_MRT          # D2Q9-only operator on a D3Q19 batch
        with pytest.raises(RuntimeError):
            b.run(bad)


def test_fp_auto_taylor_green_full_map_agreement(gpu):
    """The full 400 x 400 Taylor-Green E-RLBM map with FP32-shifted storage
    where tau - 1/2 >= 1e-3 and FP64 below (ERLBM_FP_AUTO) classifies at least
    99.9% of the points as the FP64 map does (measured 0.99952)."""
    out = {}
    for name, prec in (("fp64", gpu.FP64), ("auto", gpu.FP_AUTO)):
        jobs = gpu.tg_map_jobs(400, 400, O.COL_EXT_REG, 0.65, prec)
        with gpu.tg_batch() as b:
            out[name] = b.run(jobs)
    agree = float(np.mean(out["auto"]["unstable"] == out["fp64"]["unstable"]))
    n32 = int(np.sum(out["auto"]["precision"] != gpu.FP64))
    print(f"TG E-RLBM full map, FP_AUTO vs FP64: agreement {agree:.5f}, {n32} points in FP32 storage", flush=True)
    assert n32 > 70000
    assert agree >= 0.999
