"""Pins of the oracle's body-force path (the acc_x, acc_y, acc_z arguments of
the collision operators) -- CPU only.

  * fixtures of the reference's own step loop under a body force, compiled
    from its sources (tests/golden/ref_forced.npz, make_golden.py --only
    forced): bitwise the strict build, <= 1e-10 the FMA build;
  * closed forms that a dropped or misplaced forcing term breaks:
      - mass is conserved and, on a periodic lattice, total momentum grows by
        g a M per step, M the total mass, with g fixed by the operator:
        BGK and MRT 1; E-RLBM 1 + (1/tau_nh - 1/tau)/2 (the equilibrium at
        u + a/2 moves F/2 into f_gt1, which relaxes with tau_nh; S adds
        (1 - 1/(2 tau)) F); RLBM the same at tau_nh = 1;
      - a BGK channel driven by a body force between halfway bounce-back
        walls settles to the Poiseuille parabola; at tau - 1/2 = sqrt(3/16)
        the walls sit exactly half way, so the profile is
        u(y) = a/(2 nu) (y - 1/2)(ny - 3/2 - y) with u = m/rho + a/2.
"""
import math

import numpy as np
import pytest

import oracle as O


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_forced_trajectories_vs_fixture(orc, golden):
    g = golden("ref_forced.npz")
    # D2Q9 channel between walls, every operator
    ini, st, nf = orc.stream_map(9, 16, 12, 1, g["pois9_mask"])
    acc = tuple(g["pois9_acc"])
    for col in (O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG):
        out, n, v = orc.run_bc_steps(9, g["pois9_in"], 16, 12, 1, ini, st, 400, col, 0.8, 0.65, acc=acc)
        assert n == 400 and v == 0
        assert np.array_equal(out, g[f"pois9_c{col}_out_strict"]), col
        assert _rel(out, g[f"pois9_c{col}_out_fma"]) < 1e-10, col
    # D2Q9 periodic double shear layer
    ini, st, nf = orc.stream_map(9, 32, 32, 1)
    acc = tuple(g["dsl32f_acc"])
    for col in (O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG):
        out = orc.run_steps(9, g["dsl32f_in"], st, 200, col, 0.52, 1.0 if col == O.COL_REG else 0.65, acc=acc)
        assert np.array_equal(out, g[f"dsl32f_c{col}_out_strict"]), col
        assert _rel(out, g[f"dsl32f_c{col}_out_fma"]) < 1e-10, col
    # D3Q19 duct
    ini, st, nf = orc.stream_map(19, 6, 8, 8, g["duct19_mask"])
    acc = tuple(g["duct19_acc"])
    for col in (O.COL_BGK, O.COL_REG, O.COL_EXT_REG):
        out, n, v = orc.run_bc_steps(19, g["duct19_in"], 6, 8, 8, ini, st, 300, col, 0.7, 0.65, acc=acc)
        assert np.array_equal(out, g[f"duct19_c{col}_out_strict"]), col
        assert _rel(out, g[f"duct19_c{col}_out_fma"]) < 1e-10, col
    # D3Q19 periodic
    ini, st, nf = orc.stream_map(19, 8, 8, 8)
    out = orc.run_steps(19, g["per19f_in"], st, 100, O.COL_EXT_REG, 0.6, 0.65, acc=tuple(g["per19f_acc"]))
    assert np.array_equal(out, g[f"per19f_c{O.COL_EXT_REG}_out_strict"])
    assert _rel(out, g[f"per19f_c{O.COL_EXT_REG}_out_fma"]) < 1e-10


def _momentum_gain(col, tau, tau_nh):
    if col in (O.COL_BGK, O.COL_MRT):
        return 1.0
    if col == O.COL_REG:
        tau_nh = 1.0
    return 1.0 + 0.5 * (1.0 / tau_nh - 1.0 / tau)


@pytest.mark.parametrize("q,col", [(9, O.COL_BGK), (9, O.COL_MRT), (9, O.COL_REG), (9, O.COL_EXT_REG),
                                   (19, O.COL_BGK), (19, O.COL_REG), (19, O.COL_EXT_REG)])
def test_forced_momentum_balance(orc, q, col):
    """Periodic lattice: sum_sites sum_i c_i f grows by g a M per step and
    sum f stays M, with the operator's g (module docstring)."""
    tau, tnh, n = 0.58, 0.7, 60
    if q == 9:
        dims, acc = (24, 20, 1), np.array([3e-5, -2e-5, 0.0])
        f0 = orc.init_double_shear_d2q9(24)[: 24 * 20]
    else:
        dims, acc = (8, 6, 5), np.array([3e-5, -2e-5, 1e-5])
        rng = np.random.default_rng(7)
        f0 = np.tile(orc.dist_eq(19, 0.01, 0.02, -0.01, 1.0), (8 * 6 * 5, 1))
        f0 *= 1.0 + 1e-3 * rng.uniform(-1.0, 1.0, f0.shape)
    c = orc.lattice(q)[0]
    ini, st, nf = orc.stream_map(q, *dims)
    out = orc.run_steps(q, f0, st, n, col, tau, tnh, acc=tuple(acc))
    mass0, mass = f0.sum(), out.sum()
    assert abs(mass - mass0) / mass0 < 1e-13
    d = c.shape[1]
    dp = (out @ c).sum(axis=0) - (f0 @ c).sum(axis=0)
    want = n * _momentum_gain(col, tau, tnh) * acc[:d] * mass0
    np.testing.assert_allclose(dp, want, rtol=1e-9, atol=1e-13)
    # and a plausible slip -- the equilibrium at u instead of u + a/2 --
    # would change g by (1 - 1/tau_nh)/2 for E-RLBM: the test tells them apart
    if col == O.COL_EXT_REG:
        slip = n * (_momentum_gain(col, tau, tnh) - 0.5 * (1.0 / tnh)) * acc[:d] * mass0
        assert not np.allclose(dp, slip, rtol=1e-6, atol=0)


def test_forced_poiseuille_bgk_closed_form(orc):
    nx, ny = 8, 12
    m = np.ones(nx * ny, np.int32)
    m[[x + y * nx for y in (0, ny - 1) for x in range(nx)]] = 0
    ini, st, nf = orc.stream_map(9, nx, ny, 1, m)
    tau = 0.5 + math.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    a = 1e-6
    f = np.array([orc.dist_eq(9, 0, 0, 0, 1.0)] * nf)
    f, n, v = orc.run_bc_steps(9, f, nx, ny, 1, ini, st, 4000, O.COL_BGK, tau, tau, acc=(a, 0.0, 0.0))
    c = orc.lattice(9)[0]
    rho = f.sum(axis=1)
    ux = (f @ c)[:, 0] / rho + 0.5 * a
    uy = (f @ c)[:, 1] / rho
    ys = np.array([y for y in range(1, ny - 1) for _ in range(nx)])
    want = a / (2.0 * nu) * (ys - 0.5) * (ny - 1.5 - ys)
    np.testing.assert_allclose(ux, want, rtol=1e-8)
    assert np.abs(uy).max() < 1e-14   # round-off of sums of O(1) populations
