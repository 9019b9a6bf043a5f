"""Moment-space collision arithmetic (erlbm_set_arith(h, ERLBM_ARITH_MOMENT)).

The same BGK / RLBM / E-RLBM operators as coll_BGK, coll_Reg and coll_Ext_Reg
(SimBoltz_Functions/Collision_operators.cpp:11-49, 859-967, 1181-1307),
re-associated in moment space.  The bar is the reference's tolerance: a
maximum relative population error of at most 1e-10 after 1000 steps, against
the strict oracle (the reference's operation order) and, for the headline
workload, the FMA-contracted oracle (the reference's default build).  Each
check also asserts that the result is not bitwise the strict one, so that the
moment kernels, not the exact ones, ran.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
N = 1000


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _run(gpu, q, dims, f0, col, tau, tnh, calls, arith, fuse=None, monkeypatch=None):
    if fuse is not None:
        monkeypatch.setenv("ERLBM_FUSE", str(fuse))
    with gpu.Lattice(q, *dims) as lat:
        lat.set_params(tau, tnh, 1.0, col)
        lat.set_arith(arith)
        lat.upload(f0)
        shape = lat.step_shape()
        for n in calls:
            lat.step(n)
        assert lat.status() == (sum(calls), 0)
        return lat.download(), shape


# Taylor-Green L = 256 with f_eq + f1 initialisation (non-equilibrium
# populations from the first step), tau = 0.62: every operator
TG_L = 256
CASES = {O.COL_BGK: (0.62, 0.62), O.COL_REG: (0.62, 1.0), O.COL_EXT_REG: (0.62, 0.65)}


@pytest.fixture(scope="module")
def tg_oracle(orc):
    f0 = orc.init_taylor_green_d2q9(TG_L, 2.0 / TG_L, 0.62, True, True)
    ini, st, nf = orc.stream_map(9, TG_L, TG_L, 1)
    out = {}
    for col, (tau, tnh) in CASES.items():
        print(f"oracle: TG L={TG_L} col={col} {N} steps", flush=True)
        out[col] = orc.run_steps(9, f0, st, N, col, tau, tnh, nthreads=16)
    return f0, out


@pytest.mark.parametrize("col", sorted(CASES))
@pytest.mark.parametrize("shape", [42, 40, 26, 39, 11, 46, 47, 48, 1, 3, 7])
def test_moment_d2q9_walks_within_tolerance(gpu, monkeypatch, tg_oracle, col, shape):
    """The three-step walks (lane-aligned with 4 and 2 wavefronts, 256-wide,
    any-width), the two-step strip walk, the four-, five- and six-step
    lane-aligned walks and the 64x8, 32x8 and 64x4 two-step tiles, every
    operator."""
    f0, want = tg_oracle
    tau, tnh = CASES[col]
    print(f"moment TG L={TG_L} col={col} shape={shape}", flush=True)
    got, (f, spl, desc) = _run(gpu, 9, (TG_L, TG_L, 1), f0, col, tau, tnh, [N], gpu.ARITH_MOMENT, shape, monkeypatch)
    assert f == shape, desc
    err = _rel(got, want[col])
    print(f"  {desc}: max rel err vs strict oracle {err:.2e}", flush=True)
    assert 0.0 < err <= TOL


def test_moment_call_patterns_and_remainders(gpu, monkeypatch, tg_oracle):
    """Calls whose lengths leave remainders after the three-step launches
    (the one-step kernel, in the reference's order, joins the moment-space
    launches; the two-step remainder walk runs at L = 2048, in
    test_gpu_fullsize.py)."""
    f0, want = tg_oracle
    tau, tnh = CASES[O.COL_EXT_REG]
    got, _ = _run(gpu, 9, (TG_L, TG_L, 1), f0, O.COL_EXT_REG, tau, tnh, [500, 1, 250, 249], gpu.ARITH_MOMENT, 42,
                  monkeypatch)
    err = _rel(got, want[O.COL_EXT_REG])
    print(f"moment, calls 500+1+250+249: max rel err {err:.2e}", flush=True)
    assert 0.0 < err <= TOL


def test_exact_arith_stays_bitwise(gpu, tg_oracle):
    """Setting ERLBM_ARITH_EXACT back gives the reference's bits again."""
    f0, want = tg_oracle
    tau, tnh = CASES[O.COL_EXT_REG]
    with gpu.Lattice(9, TG_L, TG_L, 1) as lat:
        lat.set_params(tau, tnh, 1.0, O.COL_EXT_REG)
        lat.set_arith(gpu.ARITH_MOMENT)
        lat.set_arith(gpu.ARITH_EXACT)
        lat.upload(f0)
        lat.step(N)
        assert np.array_equal(lat.download(), want[O.COL_EXT_REG])


def test_moment_mrt_keeps_the_reference_order(gpu, orc):
    """coll_MRT has no moment-space form here: its lattices stay bitwise."""
    L = 128
    f0 = orc.init_double_shear_d2q9(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    tau = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
    want = orc.run_steps(9, f0, st, 200, O.COL_MRT, tau, 0.65)
    got, _ = _run(gpu, 9, (L, L, 1), f0, O.COL_MRT, tau, 0.65, [200], gpu.ARITH_MOMENT)
    assert np.array_equal(got, want)


def _smooth_d3q19(nx, ny, nz, u0=0.04):
    """Equilibrium populations of a smooth periodic 3-D velocity field
    (rho = 1 + small wave), from the second-order equilibrium."""
    c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                  [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1], [0, -1, -1],
                  [0, 1, 1], [0, -1, 1], [0, 1, -1]], float)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ux = u0 * np.sin(2 * np.pi * z / nz) * np.cos(2 * np.pi * y / ny)
    uy = u0 * np.cos(2 * np.pi * x / nx)
    uz = u0 * np.sin(2 * np.pi * y / ny) * np.sin(2 * np.pi * x / nx)
    rho = 1.0 + 0.01 * np.cos(2 * np.pi * x / nx) * np.cos(2 * np.pi * z / nz)
    u = np.stack([ux, uy, uz], -1).reshape(-1, 3)
    cu = u @ c.T
    uu = (u * u).sum(1)[:, None]
    return (w[None, :] * rho.reshape(-1, 1) * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * uu)).copy()


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_REG, O.COL_EXT_REG])
def test_moment_d3q19_zwalk_within_tolerance(gpu, orc, monkeypatch, col):
    """D3Q19 two-step z-walk (64x8 tiles), ragged segments, every operator."""
    nx, ny, nz = 128, 16, 35
    f0 = _smooth_d3q19(nx, ny, nz)
    ini, st, nf = orc.stream_map(19, nx, ny, nz)
    tau, tnh = {O.COL_BGK: (0.6, 0.6), O.COL_REG: (0.6, 1.0), O.COL_EXT_REG: (0.6, 0.65)}[col]
    print(f"oracle: D3Q19 {nx}x{ny}x{nz} col={col} {N} steps", flush=True)
    want = orc.run_steps(19, f0, st, N, col, tau, tnh, nthreads=16)
    got, (f, spl, desc) = _run(gpu, 19, (nx, ny, nz), f0, col, tau, tnh, [N], gpu.ARITH_MOMENT, 33, monkeypatch)
    assert (f, spl) == (33, 2), desc
    err = _rel(got, want)
    print(f"  {desc}: max rel err vs strict oracle {err:.2e}", flush=True)
    assert 0.0 < err <= TOL


def test_set_arith_rejects_bad_values(gpu):
    with gpu.Lattice(9, 64, 64, 1) as lat:
        with pytest.raises(gpu.ErlbmError):
            lat.set_arith(7)
    for prec in (gpu.FP16_SHIFTED, gpu.FP32_SHIFTED_F32):
        with gpu.Lattice(9, 64, 64, 1, precision=prec) as lat:
            with pytest.raises(gpu.ErlbmError):
                lat.set_arith(gpu.ARITH_MOMENT)
    with gpu.Lattice(9, 64, 64, 1, precision=gpu.FP32_SHIFTED) as lat:
        lat.set_arith(gpu.ARITH_MOMENT)
