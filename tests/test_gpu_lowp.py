"""Shifted low-precision storage of the plain lattice path (FP32 / FP16
populations f_i - w_i rho0, FP32 arithmetic, FP64 reductions) against the
FP64 oracle.  Tolerances are stated per observable, as the error budget of
the storage format, not bitwise."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _vel(f, q=9):
    c = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [1, -1], [-1, 1]], float)
    rho = f.sum(1)
    return (f @ c) / rho[:, None], rho


def _l2(u, ref):
    return float(np.sqrt(np.sum((u - ref) ** 2) / np.sum(ref ** 2)))


@pytest.mark.parametrize("prec,tol", [("FP32_SHIFTED", 1e-6), ("FP32_SHIFTED_F32", 1e-5), ("FP16_SHIFTED", 5e-2)])
def test_double_shear_shifted_vs_fp64(gpu, orc, prec, tol):
    """D2Q9 double shear layer L=256, 1000 steps: relative L2 velocity error
    against the FP64 oracle below tol (FP32 storage: 1e-6, measured 2.1e-7;
    FP32 arithmetic: 1e-5, measured 2.95e-6; FP16: 5e-2)."""
    L, n = 256, 1000
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65, nthreads=8)
    lat = gpu.Lattice(9, L, L, 1, precision=getattr(gpu, prec))
    with lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(n)
        got = lat.download()
        assert lat.status() == (n, 0)
    u, rho = _vel(got)
    uw, rhow = _vel(want)
    err = _l2(u, uw)
    print(f"{prec}: DSL L={L} {n} steps, L2 velocity error vs FP64 {err:.3e}, "
          f"mass drift {abs(rho.sum() - rhow.sum()) / rhow.sum():.2e}", flush=True)
    assert np.isfinite(got).all()
    assert err < tol
    assert abs(rho.sum() - rhow.sum()) / rhow.sum() < (1e-7 if prec.startswith("FP32") else 1e-3)


@pytest.mark.parametrize("prec,tol", [("FP32_SHIFTED", 0.01), ("FP16_SHIFTED", None), ("FP32_SHIFTED_F32", 0.01)])
def test_taylor_green_shifted_error(gpu, orc, prec, tol):
    """Config: TG D2Q9 L=64, E-RLBM 0.65, t=T.  FP32-shifted storage, with
    FP64 arithmetic or with FP32 arithmetic, keeps E_vector within 1% of the
    FP64 value 0.000298896; the FP16 variant is measured and reported."""
    L = 64
    u0 = 2.0 / L
    visc = u0 * L / 50.0
    tau = 0.5 + 3.0 * visc
    nT = int(round(L / u0))
    f0 = gpu.init_taylor_green(9, L, u0, tau)
    lat = gpu.Lattice(9, L, L, 1, precision=getattr(gpu, prec))
    with lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(nT + 1)
        post = lat.download(gpu.POST_COLLISION)
    ev, em, _ = orc.tg_error_d2q9(post, L, u0, visc, nT)
    rel = abs(ev - 0.000298896) / 0.000298896
    print(f"{prec} TG L=64 E_vector {ev:.6g} (FP64 0.000298896, {100 * rel:.2f}% off)", flush=True)
    assert np.isfinite(ev)
    if tol is not None:
        assert rel < tol


def test_obstacle_fp32_runs_like_fp64(gpu, golden):
    g = golden("ref_trajectories.npz")
    tau, tnh, nst = g["obst_stable_params"]
    lat = gpu.Lattice(19, 22, 12, 1, mask=gpu.obstacle_mask(), precision=gpu.FP32_SHIFTED)
    with lat:
        lat.set_params(tau, tnh, 1.0, O.COL_EXT_REG)
        lat.set_zouhe_inlet(1, 0.1 / np.sqrt(3.0))
        lat.set_outlet_rescale(20, 1.0)
        lat.set_density_check(10.0)
        lat.upload(g["obst_in"])
        lat.step(int(nst))
        got = lat.download()
        assert lat.status() == (int(nst), 0)
    ref = g["obst_stable_out_strict"]
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"obstacle FP32 vs FP64 after {int(nst)} steps: max abs diff / max {rel:.2e}", flush=True)
    assert rel < 1e-4


def test_shifted_energy_monitor_fp64(gpu, orc):
    """Monitors of shifted storage reduce in FP64: the energy of an uploaded
    state equals the FP64 energy of its rounded populations."""
    f0 = gpu.init_tg_map(16, 0.3)
    lat = gpu.Lattice(19, 16, 16, 1, precision=gpu.FP32_SHIFTED)
    with lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        back = lat.download()
        e = lat.kinetic_energy(gpu.STATE, ordered=True)
    assert np.max(np.abs(back - f0)) < 1e-8
    assert e == orc.kinetic_energy(19, back)


@pytest.mark.parametrize("prec,shape,remap", [("FP32_SHIFTED", 11, 1), ("FP32_SHIFTED", 13, 0),
                                              ("FP16_SHIFTED", 11, 1), ("FP16_SHIFTED", 13, 0),
                                              ("FP32_SHIFTED", 22, 1), ("FP32_SHIFTED", 24, 0),
                                              ("FP32_SHIFTED", 42, 1), ("FP32_SHIFTED", 42, 0)])
def test_two_step_strip_equals_one_step(gpu, monkeypatch, prec, shape, remap):
    """The two- and three-step strip walks keep the intermediate populations
    in LDS in the storage type, so their results equal the one-step shifted
    path's bitwise (any call pattern, monitors included)."""
    L = 256
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    p = getattr(gpu, prec)
    monkeypatch.setenv("ERLBM_FUSE", "0")
    ref = gpu.Lattice(9, L, L, 1, precision=p)
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    monkeypatch.setenv("ERLBM_FUSE_REMAP", str(remap))
    fused = gpu.Lattice(9, L, L, 1, precision=p)
    with ref, fused:
        for lat in (ref, fused):
            lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
            lat.upload(f0)
        done = 0
        for n in (1, 8, 3, 6):
            ref.step(n)
            fused.step(n)
            done += n
            assert fused.status() == ref.status() == (done, 0)
            assert np.array_equal(fused.download(), ref.download())
            assert fused.kinetic_energy(gpu.POST_COLLISION) == ref.kinetic_energy(gpu.POST_COLLISION)


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
def test_aligned_walk_fp32_every_collision(gpu, orc, monkeypatch, col):
    """The FP32 lane-aligned three-step walk for every collision model on a
    ragged lattice (1000 x 40, issue priority forced on): bitwise the
    one-step FP32 path after 1 + 3 + 7 steps."""
    nx, ny = 1000, 40
    rng = np.random.default_rng(col)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nx * ny, 9)))
    tnh = {O.COL_BGK: 0.52, O.COL_MRT: 0.65, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    monkeypatch.setenv("ERLBM_FUSE_PRIO", "1")
    out = []
    for fuse in ("0", "42"):
        monkeypatch.setenv("ERLBM_FUSE", fuse)
        with gpu.Lattice(9, nx, ny, 1, precision=gpu.FP32_SHIFTED) as lat:
            lat.set_params(0.52, tnh, 1.0, col)
            lat.upload(f0)
            for n in (1, 3, 7):
                lat.step(n)
            assert lat.status() == (11, 0)
            out.append(lat.download())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("shape,dims", [(22, (768, 64)), (24, (384, 96)), (42, (1000, 40)), (42, (2048, 33)),
                                        (42, (16, 16)), (42, (253, 17)), ("42p", (1000, 40))])
def test_three_step_fp32_non_power_of_two(gpu, orc, monkeypatch, shape, dims):
    """The FP32 three-step walks on sizes that are no power of two (42: the
    lane-aligned walk on strips of any width, down to 16 x 16): bitwise the
    one-step FP32 path."""
    nx, ny = dims
    rng = np.random.default_rng(nx)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nx * ny, 9)))
    out = []
    if shape == "42p":   # the aligned walk with its issue priority forced on
        shape = 42
        monkeypatch.setenv("ERLBM_FUSE_PRIO", "1")
    for fuse in ("0", str(shape)):
        monkeypatch.setenv("ERLBM_FUSE", fuse)
        with gpu.Lattice(9, nx, ny, 1, precision=gpu.FP32_SHIFTED) as lat:
            lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
            lat.upload(f0)
            lat.step(10)
            out.append(lat.download())
    assert np.array_equal(out[0], out[1])
