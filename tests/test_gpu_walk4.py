"""Four, five and six steps per launch (k_step4_astrip, ERLBM_FUSE=46;
k_step5_astrip, 47; k_step6_astrip, 48): the lane-aligned strip walk with
three to five LDS rings.  Same arithmetic as the one-step kernel,
so every population is bitwise the strict oracle, for every collision model,
on ragged widths and heights (strips of any width, segments of both walking
directions), for call lengths that leave one- to four-step remainders after
the multi-step launches; on FP32 shifted storage
(k_lp_step4_astrip) bitwise the one-step FP32 path."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SHAPE = 46
STEPS = {46: 4, 47: 5, 48: 6}


def _perturbed(orc, nx, ny, seed):
    rng = np.random.default_rng(seed)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    return feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9))), st


@pytest.mark.parametrize("shape", [46, 47, 48])
@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
@pytest.mark.parametrize("dims", [(1000, 24), (16, 16), (253, 16), (2048, 40), (272, 17), (4000, 16), (300, 77),
                                  (130, 33)])
def test_four_step_walk_ragged_bitwise(gpu, orc, monkeypatch, shape, col, dims):
    nx, ny = dims
    f0, st = _perturbed(orc, nx, ny, nx * ny + col)
    tau, tnh = 0.52, {O.COL_BGK: 0.52, O.COL_MRT: 0.65, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    print(f"{STEPS[shape]}-step walk dims={dims} col={col}", flush=True)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(tau, tnh, 1.0, col)
        lat.upload(f0)
        f, spl, desc = lat.step_shape()
        assert (f, spl) == (shape, STEPS[shape]), desc
        done = 0
        for n in (1, 8, 4, 3, 2, 9, 10, 5):
            lat.step(n)
            done += n
            assert lat.status() == (done, 0)
            assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, done, col, tau, tnh, nthreads=4)), done


@pytest.mark.parametrize("shape", [46, 47, 48])
def test_four_step_walk_full_size_with_remainders(gpu, orc, monkeypatch, shape):
    """The double-shear workload at L = 2048: calls of 1, 4, 7 (4 + 3: the
    three-step remainder launch), 6 (4 + 2: the two-step remainder), 5, 3
    and 9, 10, 11 and 8 steps (the four- and five-step remainders after
    five- and six-step launches), bitwise the strict oracle."""
    L, calls = 2048, (1, 4, 7, 6, 5, 3, 9, 10, 11, 8)
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, sum(calls), O.COL_EXT_REG, tau, 0.65, nthreads=16)
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        assert lat.step_shape()[:2] == (shape, STEPS[shape])
        for n in calls:
            lat.step(n)
        assert lat.status() == (sum(calls), 0)
        got = lat.download()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [46, 47, 48])
@pytest.mark.parametrize("prio", ["0", "1", "2"])
def test_four_step_walk_issue_priority_bitwise(gpu, orc, monkeypatch, shape, prio):
    nx, ny = 1000, 40
    f0, st = _perturbed(orc, nx, ny, 5)
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    monkeypatch.setenv("ERLBM_FUSE_PRIO", prio)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(13)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 13, O.COL_EXT_REG, 0.51, 0.65, nthreads=16))


@pytest.mark.parametrize("shape", [46, 47, 48])
def test_four_step_walk_fma_tolerance_1000_steps(gpu, monkeypatch, shape):
    """The 1e-10 tolerance against the FMA-contracted oracle after 1000
    steps (D2Q9 double shear, L = 512)."""
    ofma = O.Oracle(fma=True)
    L = 512
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = ofma.stream_map(9, L, L, 1)
    want = ofma.run_steps(9, f0, st, 1000, O.COL_EXT_REG, tau, 0.65, nthreads=16)
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(1000)
        got = lat.download()
    err = float(np.max(np.abs(got - want) / np.abs(want)))
    print(f"four-step walk vs FMA oracle, L={L}, 1000 steps: max rel err {err:.2e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("shape", [46, 47])
@pytest.mark.parametrize("dims", [(1000, 24), (253, 16), (2048, 40), (300, 77)])
def test_four_step_walk_fp32_storage_bitwise_one_step(gpu, orc, monkeypatch, shape, dims):
    """FP32 shifted storage (FP64 arithmetic): the four- and five-step walks
    keep the FP32 values the one-step launches would store, so they are
    bitwise the one-step FP32 path, for call lengths leaving every
    remainder."""
    nx, ny = dims
    f0, _ = _perturbed(orc, nx, ny, nx + 7 * ny)
    out = []
    for fuse in (str(shape), "0"):
        monkeypatch.setenv("ERLBM_FUSE", fuse)
        with gpu.Lattice(9, nx, ny, 1, precision=gpu.FP32_SHIFTED) as lat:
            lat.set_params(0.52, 0.65, 1.0, O.COL_EXT_REG)
            lat.upload(f0)
            assert lat.step_shape()[:2] == ((shape, STEPS[shape]) if fuse != "0" else (0, 1))
            for n in (1, 8, 3, 2, 9, 10, 4):
                lat.step(n)
            out.append(lat.download())
    assert np.array_equal(out[0], out[1])
