"""erlbm_batch_run_noalloc against erlbm_batch_run: the same jobs give the
same results field for field (bitwise; the same kernels run), on calls in a
row whose second is smaller than the first, across two geometries sharing
the one device scratch array, with several (precision, collision) groups per
call; more than ERLBM_NOALLOC_MAX_JOBS jobs is an argument error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    for name in a.dtype.names:
        x, y = a[name], b[name]
        if np.issubdtype(x.dtype, np.floating):
            assert np.array_equal(x, y, equal_nan=True), name
        else:
            assert np.array_equal(x, y), name


def _tg_jobs(E, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for col, tnh, prec in ((E.COL_EXT_REG, 0.65, E.FP64), (E.COL_REG, 1.0, E.FP64),
                           (E.COL_EXT_REG, 0.65, E.FP32_SHIFTED), (E.COL_EXT_REG, 0.65, E.FP_AUTO)):
        jobs = E.tg_map_jobs(400, 400, col, tnh, prec)
        out.append(jobs[rng.choice(jobs.size, n // 4, replace=False)])
    jobs = np.concatenate(out)
    return jobs[rng.permutation(jobs.size)]


def test_noalloc_batch_equals_batch_run(gpu):
    E = gpu
    with E.tg_batch() as b:
        for n, seed in ((3000, 1), (400, 2)):     # the second call smaller than the first
            jobs = _tg_jobs(E, n, seed)
            print(f"noalloc: Taylor-Green batch of {jobs.size} jobs (seed {seed})", flush=True)
            want = b.run(jobs)
            got = b.run_noalloc(jobs)
            _same(got, want)
            assert got["done"].all()
    with E.obstacle_batch() as b:
        jobs = E.obstacle_map_jobs(400, 400, 0.1, E.FP64)
        jobs = jobs[np.random.default_rng(3).choice(jobs.size, 200, replace=False)]
        print(f"noalloc: obstacle batch of {jobs.size} jobs", flush=True)
        _same(b.run_noalloc(jobs), b.run(jobs))


def test_noalloc_batch_limits(gpu):
    E = gpu
    from erlbm_amd._lib import NOALLOC_MAX_JOBS
    with E.tg_batch() as b:
        jobs = E.tg_map_jobs(400, 400, E.COL_EXT_REG, 0.65, E.FP64, i0=0, i1=400, j0=0, j1=41)
        assert jobs.size > NOALLOC_MAX_JOBS
        with pytest.raises(E.ErlbmError) as e:
            b.run_noalloc(jobs)
        assert e.value.code == 1
        assert b.run_noalloc(jobs[:0]).size == 0
        # exactly the maximum, budgets cut short so the call is quick
        full = jobs[:NOALLOC_MAX_JOBS].copy()
        full["n_steps"] = np.minimum(full["n_steps"], 300)
        print(f"noalloc: {full.size} jobs (the maximum), 300 steps each", flush=True)
        _same(b.run_noalloc(full), b.run(full))
