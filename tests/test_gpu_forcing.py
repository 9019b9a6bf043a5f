"""Body force on the GPU path (erlbm_set_body_force): the collision
operators' acc_x, acc_y, acc_z arguments (Collision_operators.cpp:1182,
1196-1254; source(), LBM_functions.cpp:5645-5716).

Bar: bitwise the reference's strict build (fixtures written from its own
sources) and the strict oracle; the closed-form momentum balance at a size the
oracle would take long over.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TAU, TAU_NH = 0.62, 0.65   # the per-site fixtures' parameters


@pytest.mark.parametrize("q,col", [(9, O.COL_BGK), (9, O.COL_MRT), (9, O.COL_REG), (9, O.COL_EXT_REG),
                                   (19, O.COL_BGK), (19, O.COL_REG), (19, O.COL_EXT_REG)])
def test_forced_collision_bitwise_vs_reference_fixture(gpu, golden, q, col):
    """One collision of 64 seeded non-equilibrium sites under a body force:
    the first step after an upload collides in place, so the post-collision
    download is colisao_sitio's output."""
    g = golden("ref_functions.npz")
    acc = (1e-4, -2e-4, 3e-5 if q == 19 else 0.0)
    print(f"forced collision q={q} col={col} acc={acc}", flush=True)
    with gpu.Lattice(q, 8, 8, 1) as lat:
        lat.set_params(TAU, TAU_NH, 1.0, col)
        lat.set_body_force(*acc)
        lat.upload(g[f"sites_q{q}"])
        lat.step(1)
        got = lat.download(gpu.POST_COLLISION)
    assert np.array_equal(got, g[f"coll_q{q}_c{col}_forced_strict"])


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
def test_forced_channel_d2q9_bitwise(gpu, orc, golden, col):
    """Body-force-driven channel between halfway bounce-back walls, 400 steps."""
    g = golden("ref_forced.npz")
    print(f"forced d2q9 channel col={col}", flush=True)
    with gpu.Lattice(9, 16, 12, 1, mask=g["pois9_mask"]) as lat:
        lat.set_params(0.8, 0.65, 1.0, col)
        lat.set_body_force(*g["pois9_acc"])
        lat.upload(g["pois9_in"])
        lat.step(400)
        assert lat.status() == (400, 0)
        got = lat.download()
    assert np.array_equal(got, g[f"pois9_c{col}_out_strict"])


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
def test_forced_periodic_d2q9_bitwise(gpu, golden, col):
    g = golden("ref_forced.npz")
    print(f"forced d2q9 periodic col={col}", flush=True)
    with gpu.Lattice(9, 32, 32, 1) as lat:
        lat.set_params(0.52, 1.0 if col == O.COL_REG else 0.65, 1.0, col)
        lat.set_body_force(*g["dsl32f_acc"])
        lat.upload(g["dsl32f_in"])
        lat.step(200)
        got = lat.download()
    assert np.array_equal(got, g[f"dsl32f_c{col}_out_strict"])


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_REG, O.COL_EXT_REG])
def test_forced_duct_d3q19_bitwise(gpu, golden, col):
    g = golden("ref_forced.npz")
    print(f"forced d3q19 duct col={col}", flush=True)
    with gpu.Lattice(19, 6, 8, 8, mask=g["duct19_mask"]) as lat:
        lat.set_params(0.7, 0.65, 1.0, col)
        lat.set_body_force(*g["duct19_acc"])
        lat.upload(g["duct19_in"])
        lat.step(300)
        got = lat.download()
    assert np.array_equal(got, g[f"duct19_c{col}_out_strict"])


def test_forced_periodic_d3q19_bitwise(gpu, golden):
    g = golden("ref_forced.npz")
    print("forced d3q19 periodic 8^3", flush=True)
    with gpu.Lattice(19, 8, 8, 8) as lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        lat.set_body_force(*g["per19f_acc"])
        lat.upload(g["per19f_in"])
        lat.step(100)
        got = lat.download()
    assert np.array_equal(got, g[f"per19f_c{O.COL_EXT_REG}_out_strict"])


def test_forced_dsl_bitwise_vs_oracle_and_momentum_balance(gpu, orc):
    """D2Q9 double shear layer L = 256 under a body force, 300 steps: bitwise
    the strict oracle; total momentum grows by (1 + (1/tau_nh - 1/tau)/2) a M
    per step (closed form of the E-RLBM forcing) at L = 1024 as well."""
    acc = (2e-6, -1e-6, 0.0)
    for L, n, bitwise in ((256, 300, True), (1024, 500, False)):
        print(f"forced dsl L={L} steps={n}", flush=True)
        tau = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
        f0 = gpu.init_double_shear(L)
        with gpu.Lattice(9, L, L, 1) as lat:
            lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
            lat.set_body_force(*acc)
            lat.upload(f0)
            assert lat.step_shape()[0] == 0   # one step per launch under a force
            lat.step(n)
            got = lat.download()
        if bitwise:
            ini, st, nf = orc.stream_map(9, L, L, 1)
            want = orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65, nthreads=16, acc=acc)
            assert np.array_equal(got, want)
        c = orc.lattice(9)[0]
        mass0 = f0.sum()
        dp = (got @ c).sum(axis=0) - (f0 @ c).sum(axis=0)
        gain = 1.0 + 0.5 * (1.0 / 0.65 - 1.0 / tau)
        np.testing.assert_allclose(dp, n * gain * np.array(acc[:2]) * mass0, rtol=1e-8)
        assert abs(got.sum() - mass0) / mass0 < 1e-12


def test_zero_force_is_the_unforced_path(gpu, orc):
    """Setting a force and clearing it again leaves the multi-step default
    kernels in charge, bitwise the unforced oracle."""
    L, n = 256, 31
    tau = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
    f0 = gpu.init_double_shear(L)
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.set_body_force(1e-5, 0.0)
        lat.set_body_force(0.0, 0.0, 0.0)
        lat.upload(f0)
        lat.step(n)
        got = lat.download()
    ini, st, nf = orc.stream_map(9, L, L, 1)
    assert np.array_equal(got, orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65, nthreads=16))


def test_body_force_argument_errors(gpu):
    with gpu.Lattice(9, 16, 16, 1, precision=gpu.FP32_SHIFTED) as lat:
        with pytest.raises(gpu.ErlbmError) as e:
            lat.set_body_force(1e-5, 0.0)
        assert e.value.code == 5
        lat.set_body_force(0.0, 0.0)   # zero force is the unforced operator, allowed
    with gpu.Lattice(9, 16, 16, 1) as lat:
        with pytest.raises(gpu.ErlbmError) as e:
            lat.set_body_force(float("nan"), 0.0)
        assert e.value.code == 1
        lat.set_body_force(0.0, 0.0, 1e-3)   # D2Q9 ignores az: accepted, no force
