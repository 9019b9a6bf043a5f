"""The moment-space form of the collision operators (the formula of
erlbm_amd/csrc/collide_moment.hpp, restated here in numpy) against the
oracle's collision in the reference's operation order
(SimBoltz_Functions/Collision_operators.cpp:11-49 coll_BGK, :859-967
coll_Reg, :1181-1307 coll_Ext_Reg), on the CPU:

    f*_i = W f_i + w_i (alpha + beta.c_i + c_i.Gamma.c_i)
    alpha    = (K0 + D K1/2) rho - 1.5 T (j.u) - 1.5 K1 tr Pi
    beta     = 3 K0 j
    Gamma_ab = 4.5 T j_a u_b + 4.5 K1 Pi_ab - 1.5 K1 rho d_ab

with (W, K0, K1) = (1 - 1/tau_nh, 1/tau_nh, 1/tau_nh - 1/tau) for E-RLBM,
(0, 1, 1 - 1/tau) for RLBM, (1 - 1/tau, 1/tau, 0) for BGK and T = 1/tau.
On non-equilibrium populations (random perturbations of an equilibrium,
so every moment of the non-equilibrium part is present) it must equal the
reference operator to rounding, for D2Q9 and D3Q19; each of a few
plausible mistakes (a dropped stress term, the trace term, a swapped
relaxation time, the equilibrium at the wrong velocity) must not.
"""
import numpy as np
import pytest

import oracle as O

C9 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0], [-1, -1, 0], [1, -1, 0],
               [-1, 1, 0]], float)
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1], [0, -1, -1],
                [0, 1, 1], [0, -1, 1], [0, 1, -1]], float)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _consts(col, tau, tnh):
    if col == O.COL_BGK:
        return 1 - 1 / tau, 1 / tau, 0.0
    if col == O.COL_REG:
        return 0.0, 1.0, 1 - 1 / tau
    return 1 - 1 / tnh, 1 / tnh, 1 / tnh - 1 / tau


def moment_collide(f, q, col, tau, tnh, mutate=None):
    c, w = (C9, W9) if q == 9 else (C19, W19)
    d = 2 if q == 9 else 3
    c = c[:, :d]
    W, K0, K1 = _consts(col, tau, tnh)
    if mutate == "swap_tau":
        W, K0, K1 = _consts(col, tnh, tau)
    T = 1 / tau
    rho = f.sum(1)
    j = f @ c
    u = j / rho[:, None]
    pi = np.einsum("ni,ia,ib->nab", f, c, c)
    ju = j[:, :, None] * u[:, None, :]
    if mutate == "u_shift":
        ju = j[:, :, None] * (u[:, None, :] * 1.01)
    trp = np.trace(pi, axis1=1, axis2=2)
    kt = 0.0 if mutate == "no_trace" else 1.5 * K1 * trp
    alpha = (K0 + d * K1 / 2) * rho - 1.5 * T * np.einsum("na,na->n", j, u) - kt
    beta = 3 * K0 * j
    k45 = 0.0 if mutate == "no_stress" else 4.5 * K1
    gamma = 4.5 * T * ju + k45 * pi - 1.5 * K1 * rho[:, None, None] * np.eye(d)[None]
    a = alpha[:, None] + beta @ c.T + np.einsum("ia,nab,ib->ni", c, gamma, c)
    return W * f + w[None, :] * a


def _states(orc, q, n, seed):
    rng = np.random.default_rng(seed)
    feq = np.stack([orc.dist_eq(q, *rng.uniform(-0.08, 0.08, 3) * ([1, 1, 0] if q == 9 else 1), 1.0 + 0.05 * rng.uniform(-1, 1))
                    for _ in range(n)])
    return feq * (1 + 0.05 * rng.uniform(-1, 1, feq.shape))


CASES = [(O.COL_BGK, 0.62, 0.62), (O.COL_REG, 0.62, 1.0), (O.COL_EXT_REG, 0.62, 0.65), (O.COL_EXT_REG, 0.5024, 0.65),
         (O.COL_EXT_REG, 0.9, 0.55)]


@pytest.mark.parametrize("q", [9, 19])
@pytest.mark.parametrize("col,tau,tnh", CASES)
def test_moment_form_equals_reference_operator(orc, q, col, tau, tnh):
    f = _states(orc, q, 400, q * 100 + col)
    want = orc.collide(f, col, tau, tnh)
    got = moment_collide(f, q, col, tau, tnh)
    err = np.max(np.abs(got - want) / np.abs(want))
    assert err < 1e-13, err


@pytest.mark.parametrize("q", [9, 19])
@pytest.mark.parametrize("mutate", ["swap_tau", "no_trace", "no_stress", "u_shift"])
def test_moment_form_mutations_fail(orc, q, mutate):
    col, tau, tnh = O.COL_EXT_REG, 0.62, 0.8
    f = _states(orc, q, 400, q + 7)
    want = orc.collide(f, col, tau, tnh)
    got = moment_collide(f, q, col, tau, tnh, mutate=mutate)
    assert np.max(np.abs(got - want) / np.abs(want)) > 1e-6
