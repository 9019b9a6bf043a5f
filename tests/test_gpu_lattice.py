"""Parity of the plain (one lattice per handle) HIP path against the oracle.

FP64 integer-exact bar: every population bitwise equal to the strict oracle
(the reference's operation order without FMA contraction); against the
FMA-contracted build, max relative population error <= 1e-10 after 1000
steps, as the target states.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL_FMA = 1e-10


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _periodic(gpu, q, L, tau, tnh, col, f0, nz=1):
    lat = gpu.Lattice(q, L, L, nz)
    lat.set_params(tau, tnh, 1.0, col)
    lat.upload(f0)
    return lat


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
def test_d2q9_double_shear_bitwise(gpu, orc, col):
    L, n = 64, 200
    print(f"d2q9 dsl L={L} col={col} steps={n}", flush=True)
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    tnh = {O.COL_BGK: tau, O.COL_MRT: 0.65, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, n, col, tau, tnh, nthreads=4)
    with _periodic(gpu, 9, L, tau, tnh, col, f0) as lat:
        lat.step(n)
        got = lat.download()
        assert lat.status() == (n, 0)
    assert np.array_equal(got, want)


def test_upload_download_roundtrip(gpu, orc):
    f0 = gpu.init_double_shear(48)
    with _periodic(gpu, 9, 48, 0.6, 0.65, O.COL_EXT_REG, f0) as lat:
        assert np.array_equal(lat.download(), f0)
        lat.step(1)
        post = lat.download(gpu.POST_COLLISION)
        want = orc.collide(f0, O.COL_EXT_REG, 0.6, 0.65)
        assert np.array_equal(post, want)


def test_taylor_green_published_error(gpu, orc):
    """Config: D2Q9 TG L=64, tau=0.62, E-RLBM tau_nh=0.65, f_eq+f1 init, t=T.
    data/fig3_taylor_green_Re50.csv:7 publishes E_vector 0.000298896."""
    L = 64
    u0 = 2.0 / L
    visc = u0 * L / 50.0
    tau = 0.5 + 3.0 * visc
    nT = int(round(L / u0))
    f0 = gpu.init_taylor_green(9, L, u0, tau)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        lat.step(nT + 1)
        post = lat.download(gpu.POST_COLLISION)
    ev, em, _ = orc.tg_error_d2q9(post, L, u0, visc, nT)
    assert "%.6g" % ev == "0.000298896"
    assert "%.6g" % em == "0.000298724"


@pytest.mark.parametrize("col,tnh", [(O.COL_REG, 1.0), (O.COL_EXT_REG, 0.65)])
def test_d3q19_tg_map_field_bitwise_and_energy(gpu, orc, col, tnh):
    L, n = 16, 500
    f0 = gpu.init_tg_map(L, 0.3)
    ini, st, nf = orc.stream_map(19, L, L, 1)
    want = orc.run_steps(19, f0, st, n, col, 0.5003, tnh)
    with _periodic(gpu, 19, L, 0.5003, tnh, col, f0) as lat:
        assert lat.kinetic_energy(gpu.STATE, ordered=True) == orc.kinetic_energy(19, f0)
        lat.step(n)
        got = lat.download()
        assert np.array_equal(got, want)
        assert lat.kinetic_energy(gpu.STATE, ordered=True) == orc.kinetic_energy(19, want)
        e_fast = lat.kinetic_energy(gpu.STATE, ordered=False)
        assert abs(e_fast - orc.kinetic_energy(19, want)) <= 1e-13 * abs(e_fast)


def test_obstacle_channel_bitwise(gpu, orc, golden):
    g = golden("ref_trajectories.npz")
    m = gpu.obstacle_mask()
    for name in ("stable", "unstable"):
        tau, tnh, nst = g[f"obst_{name}_params"]
        print(f"obstacle {name}", flush=True)
        with gpu.Lattice(19, 22, 12, 1, mask=m) as lat:
            lat.set_params(tau, tnh, 1.0, O.COL_EXT_REG)
            lat.set_zouhe_inlet(1, 0.1 / np.sqrt(3.0))
            lat.set_outlet_rescale(20, 1.0)
            lat.set_density_check(10.0)
            lat.upload(g["obst_in"])
            lat.step(int(nst))
            steps, reason = lat.status()
            got = lat.download()
        want_steps, want_viol = g[f"obst_{name}_steps_strict"]
        assert steps == want_steps
        assert (reason == gpu.STOP_DENSITY) == bool(want_viol)
        assert np.array_equal(got, g[f"obst_{name}_out_strict"], equal_nan=True)


def test_d2q9_channel_with_boundaries(gpu, orc, golden):
    g = golden("ref_trajectories.npz")
    with gpu.Lattice(9, 24, 10, 1, mask=g["ch9_mask"]) as lat:
        lat.set_params(0.55, 0.7, 1.0, O.COL_EXT_REG)
        lat.set_zouhe_inlet(1, 0.05)
        lat.set_outlet_rescale(22, 1.0)
        lat.set_density_check(10.0)
        lat.upload(g["ch9_in"])
        lat.step(300)
        assert lat.status() == (300, 0)
        assert np.array_equal(lat.download(), g["ch9_out_strict"])


def test_energy_check_stops_like_map_driver(gpu, orc):
    """Unstable Taylor-Green map point: same stop step as roda_ponto."""
    u0, tau, tnh = orc.tg_map_params(380, 5, col=O.COL_REG, tau_nh=1.0)
    unstable, stop, _ = orc.run_tg_map_point(16, O.COL_REG, u0, tau, tnh)
    assert unstable == 1
    n_steps = int((100 * 16) / u0)
    with _periodic(gpu, 19, 16, tau, tnh, O.COL_REG, gpu.init_tg_map(16, u0)) as lat:
        lat.step(n_steps, energy_interval=100, energy_factor=2.0)
        steps, reason = lat.status()
    assert reason == gpu.STOP_ENERGY
    assert steps == stop


def test_fma_reference_tolerance_1000_steps(gpu, golden):
    """1e-10 max relative population error against the FMA-contracted
    reference after 1000 steps (D2Q9 DSL L=128)."""
    import oracle
    ofma = oracle.Oracle(fma=True)
    L = 128
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = ofma.stream_map(9, L, L, 1)
    want = ofma.run_steps(9, f0, st, 1000, O.COL_EXT_REG, tau, 0.65, nthreads=8)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        lat.step(1000)
        got = lat.download()
    assert _rel(got, want) <= TOL_FMA


@pytest.mark.parametrize("npt,ntl,pad,remap,pair,layout", [(1, 0, 320, 1, 0, 0), (2, 0, 0, 1, 0, 1), (2, 1, 64, 0, 0, 0),
                                                          (4, 1, 4160, 1, 0, 1), (1, 1, 0, 0, 1, 0), (1, 0, 64, 1, 1, 1),
                                                          (1, 1, 0, 0, 0, 1), (1, 2, 64, 1, 0, 0)])
def test_periodic_launch_variants_bitwise(gpu, orc, monkeypatch, npt, ntl, pad, remap, pair, layout):
    """Every launch shape and layout the tuner may pick gives the same populations."""
    for k, v in (("ERLBM_NPT", npt), ("ERLBM_NT_LOADS", ntl), ("ERLBM_PLANE_PAD", pad), ("ERLBM_XCD_REMAP", remap),
                 ("ERLBM_PAIR", pair), ("ERLBM_LAYOUT", layout), ("ERLBM_VEC", 0)):
        monkeypatch.setenv(k, str(v))
    L, n = 96, 30
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        lat.step(n)
        assert np.array_equal(lat.download(), want)


@pytest.mark.parametrize("layout", [0, 1])
def test_layouts_obstacle_and_monitors(gpu, orc, golden, monkeypatch, layout):
    """Both population layouts through the general kernel, transfers and monitors."""
    monkeypatch.setenv("ERLBM_LAYOUT", str(layout))
    g = golden("ref_trajectories.npz")
    tau, tnh, nst = g["obst_stable_params"]
    with gpu.Lattice(19, 22, 12, 1, mask=gpu.obstacle_mask()) as lat:
        lat.set_params(tau, tnh, 1.0, O.COL_EXT_REG)
        lat.set_zouhe_inlet(1, 0.1 / np.sqrt(3.0))
        lat.set_outlet_rescale(20, 1.0)
        lat.set_density_check(10.0)
        lat.upload(g["obst_in"])
        lat.step(int(nst))
        got = lat.download()
        assert np.array_equal(got, g["obst_stable_out_strict"])
        assert lat.kinetic_energy(gpu.STATE, ordered=True) == orc.kinetic_energy(19, got)


@pytest.mark.parametrize("vec,layout,q,L", [(1, 1, 9, 96), (2, 1, 9, 96), (1, 0, 9, 64), (1, 1, 19, 16), (2, 1, 19, 16),
                                            (1, 1, 9, 2048)])
def test_vectorised_kernel_bitwise(gpu, orc, monkeypatch, vec, layout, q, L):
    """The 16-byte vectorised periodic kernel (shuffle or edge-load variant)
    gives the reference's populations bit for bit, D2Q9 and D3Q19."""
    monkeypatch.setenv("ERLBM_VEC", str(vec))
    monkeypatch.setenv("ERLBM_LAYOUT", str(layout))
    n = 40 if L < 1000 else 3
    if q == 9:
        tau = 0.5 + 3. * (1. / 32.) * L / 80000.
        f0 = gpu.init_double_shear(L)
        tnh = 0.65
    else:
        tau, tnh = 0.5003, 0.65
        f0 = gpu.init_tg_map(L, 0.3)
    ini, st, nf = orc.stream_map(q, L, L, 1)
    want = orc.run_steps(q, f0, st, n, O.COL_EXT_REG, tau, tnh, nthreads=8)
    with _periodic(gpu, q, L, tau, tnh, O.COL_EXT_REG, f0) as lat:
        lat.step(n)
        assert np.array_equal(lat.download(), want)


@pytest.mark.parametrize("q,L,prec", [(9, 96, "FP64"), (19, 16, "FP64"), (9, 64, "FP16_SHIFTED")])
def test_autotune_leaves_the_run_alone(gpu, orc, q, L, prec):
    """The autotuner times candidate launch shapes on a scratch copy: the run
    continues from where it was, bitwise as if no tuning happened."""
    p = getattr(gpu, prec)
    if q == 9:
        tau = 0.5 + 3. * (1. / 32.) * L / 80000.
        f0 = gpu.init_double_shear(L)
    else:
        tau = 0.503
        f0 = gpu.init_tg_map(L, 0.05)
    ini, st, nf = orc.stream_map(q, L, L, 1)
    with gpu.Lattice(q, L, L, 1, precision=p) as a, gpu.Lattice(q, L, L, 1, precision=p) as b:
        for lat in (a, b):
            lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
            lat.upload(f0)
            lat.step(7)
        name, us = a.autotune(20)
        print(f"q={q} L={L} {prec}: picked '{name}' at {us:.2f} us/step", flush=True)
        assert us > 0 and name
        a.step(13)
        b.step(13)
        assert np.array_equal(a.download(), b.download())
        if prec == "FP64":
            assert np.array_equal(a.download(), orc.run_steps(q, f0, st, 20, O.COL_EXT_REG, tau, 0.65))


@pytest.mark.parametrize("vec", [0, 1])
def test_mrt_channel_and_vector_kernel(gpu, orc, golden, monkeypatch, vec):
    """coll_MRT (D2Q9) through the boundary-column kernel (Zou-He inlet,
    outlet rescale, bounce-back) and the periodic kernels: bitwise the
    oracle's MRT, which is pinned to the reference's coll_MRT."""
    monkeypatch.setenv("ERLBM_VEC", str(vec))
    g = golden("ref_trajectories.npz")
    m = g["ch9_mask"]
    ini, st, nf = orc.stream_map(9, 24, 10, 1, m)
    want, n, v = orc.run_bc_steps(9, g["ch9_in"], 24, 10, 1, ini, st, 300, O.COL_MRT, 0.55, 0.7, inlet_x=1,
                                  u_in=0.05, outlet_x=22, rho_out=1.0, dens_limit=10.0)
    with gpu.Lattice(9, 24, 10, 1, mask=m) as lat:
        lat.set_params(0.55, 0.7, 1.0, O.COL_MRT)
        lat.set_zouhe_inlet(1, 0.05)
        lat.set_outlet_rescale(22, 1.0)
        lat.set_density_check(10.0)
        lat.upload(g["ch9_in"])
        lat.step(300)
        assert lat.status() == (n, v)
        assert np.array_equal(lat.download(), want)
    L = 96
    tau = 0.51
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, 50, O.COL_MRT, tau, 0.65)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_MRT, f0) as lat:
        lat.step(50)
        assert np.array_equal(lat.download(), want)


def test_mrt_is_d2q9_only(gpu):
    with gpu.Lattice(19, 8, 8, 2) as lat:
        with pytest.raises(RuntimeError, match="MRT"):
            lat.set_params(0.6, 0.65, 1.0, O.COL_MRT)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                   36, 37, 38, 39, 40, 41, 42, 43, 44, 45])
def test_two_step_fused_kernel_bitwise(gpu, orc, monkeypatch, shape):
    """Two or three steps per launch (tiles, strip walks, wavefront strips):
    any mix of call lengths gives the populations, post-collision monitors
    and energies of the single-step path, bitwise."""
    L = 256
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    monkeypatch.setenv("ERLBM_FUSE", "0")
    ref = _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0)
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    fused = _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0)
    with ref, fused:
        done = 0
        for n in (1, 9, 2, 7, 4):
            ref.step(n)
            fused.step(n)
            done += n
            assert ref.status() == fused.status() == (done, 0)
            assert np.array_equal(fused.download(), ref.download())
            assert np.array_equal(fused.download(gpu.POST_COLLISION), ref.download(gpu.POST_COLLISION))
            assert fused.kinetic_energy(gpu.POST_COLLISION) == ref.kinetic_energy(gpu.POST_COLLISION)
        assert np.array_equal(fused.download(), orc.run_steps(9, f0, st, done, O.COL_EXT_REG, tau, 0.65))
        # an energy-checked call falls back to single steps and stays exact
        fused.step(6, energy_interval=4, energy_factor=2.0)
        ref.step(6, energy_interval=4, energy_factor=2.0)
        fused.step(4)
        ref.step(4)
        assert fused.status() == ref.status()
        assert np.array_equal(fused.download(), ref.download())


@pytest.mark.parametrize("q,col", [(9, O.COL_EXT_REG), (9, O.COL_MRT), (9, O.COL_BGK), (19, O.COL_EXT_REG),
                                   (19, O.COL_REG)])
def test_extreme_density_scales_bitwise(gpu, orc, q, col):
    """Sites whose populations are scaled by 2^e, e in [-1000, 1000]: one
    collision equals the strict oracle's, bitwise, on both sides of the
    short reciprocal sequence's range (|rho| in [2^-600, 2^600]) and with the
    plain division beyond it."""
    rng = np.random.default_rng(5)
    dims = (16, 8, 1) if q == 9 else (8, 4, 4)
    n = dims[0] * dims[1] * dims[2]
    base = orc.dist_eq(q, 0.03, -0.02, 0.01 if q == 19 else 0.0, 1.0)
    f = base[None] * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, (n, q)))
    e = rng.integers(-700, 701, n)
    e[:10] = [-601, -600, -599, 599, 600, 601, -1000, 1000, 0, -1]
    f = np.ldexp(f, e[:, None])
    print(f"extreme density scales q={q} col={col}", flush=True)
    with gpu.Lattice(q, *dims) as lat:
        lat.set_params(0.6, 0.65, 1.0, col)
        lat.upload(f)
        lat.step(1)
        got = lat.download(gpu.POST_COLLISION)
    want = orc.collide(f, col, 0.6, 0.65)
    assert np.isfinite(want).all()
    assert np.array_equal(got, want)


def test_stop_tests_do_not_pin_single_steps(gpu, orc, monkeypatch):
    """A call with an energy test runs one step per launch; a later call
    without one goes back to the multi-step kernel when no stop was recorded
    (bitwise the oracle), and leaves a stopped run where it stopped."""
    L = 256
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    monkeypatch.setenv("ERLBM_FUSE", "42")
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        assert lat.step_shape()[:2] == (42, 3)
        lat.step(9, energy_interval=4, energy_factor=2.0)
        assert lat.status() == (9, 0)
        assert lat.step_shape()[:2] == (42, 3)
        lat.step(20)
        assert lat.status() == (29, 0)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 29, O.COL_EXT_REG, tau, 0.65))
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        lat.step(9, energy_interval=4, energy_factor=0.5)   # E / E0 ~ 1 > 0.5: stops before step 4
        assert lat.status() == (4, gpu.STOP_ENERGY)
        stopped = lat.download()
        assert np.array_equal(stopped, orc.run_steps(9, f0, st, 4, O.COL_EXT_REG, tau, 0.65))
        assert lat.step_shape()[:2] == (0, 1)
        lat.step(20)
        assert lat.status() == (4, gpu.STOP_ENERGY)
        assert np.array_equal(lat.download(), stopped)


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
@pytest.mark.parametrize("shape,dims", [(16, (300, 77)), (18, (64, 5)), (17, (130, 33)), (18, (125, 64))])
def test_wavefront_strips_ragged_bitwise(gpu, orc, monkeypatch, col, shape, dims):
    """The wavefront-strip two-step kernel on lattices whose width is no
    multiple of its 62-wide strips and whose height is no multiple of its
    segments (down to the 64-column minimum): bitwise the oracle for odd and
    even call lengths, for every collision model."""
    nx, ny = dims
    rng = np.random.default_rng(nx * ny + col)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    tau, tnh = 0.52, {O.COL_BGK: 0.52, O.COL_MRT: 0.65, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    print(f"wave strips shape={shape} dims={dims} col={col}", flush=True)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(tau, tnh, 1.0, col)
        lat.upload(f0)
        done = 0
        for n in (1, 6, 3, 10):
            lat.step(n)
            done += n
            assert lat.status() == (done, 0)
            assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, done, col, tau, tnh, nthreads=4)), done


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_MRT, O.COL_REG, O.COL_EXT_REG])
@pytest.mark.parametrize("shape,dims", [(36, (300, 77)), (37, (130, 33)), (38, (2048, 40)), (39, (1000, 24)),
                                        (38, (17, 19)), (36, (16, 16)), (39, (253, 16)), (40, (130, 33)),
                                        (41, (300, 77)), (42, (1000, 24)), (42, (16, 16)), (42, (253, 16)),
                                        (43, (1500, 40)), (43, (509, 17)), (44, (2048, 40)), (44, (317, 20)),
                                        (45, (1000, 33)), (42, (2048, 40)), (42, (4096, 16)), (42, (4000, 16)),
                                        (42, (272, 17))])
def test_any_width_strips_ragged_bitwise(gpu, orc, monkeypatch, col, shape, dims):
    """The three-step walk on strips of any width (widths split evenly, at
    most 64 NW - 4 for NW wavefronts; 40-43: lanes aligned across the
    phases, c_x = 0 populations in registers) on lattices of ragged widths
    and heights, down to 16 x 16 (one strip as wide as the lattice): bitwise
    the oracle for call lengths that are and are not multiples of three,
    for every collision model."""
    nx, ny = dims
    rng = np.random.default_rng(nx * ny + col)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    tau, tnh = 0.52, {O.COL_BGK: 0.52, O.COL_MRT: 0.65, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    print(f"any-width strips shape={shape} dims={dims} col={col}", flush=True)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(tau, tnh, 1.0, col)
        lat.upload(f0)
        done = 0
        for n in (1, 6, 4, 9):
            lat.step(n)
            done += n
            assert lat.status() == (done, 0)
            assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, done, col, tau, tnh, nthreads=4)), done


@pytest.mark.parametrize("shape,dims", [(13, (384, 96)), (11, (768, 64)), (12, (768, 48)), (1, (192, 40)), (16, (1000, 24)),
                                        (20, (384, 80)), (19, (768, 40)), (22, (768, 64)), (24, (384, 96)), (25, (1536, 64)),
                                        (26, (768, 83)), (27, (384, 50)), (29, (512, 16)), (3, (480, 40)),
                                        (3, (96, 24))])
def test_two_step_kernels_non_power_of_two(gpu, orc, monkeypatch, shape, dims):
    """The two-step kernels on widths and heights that are no power of two
    (row and plane offsets then share bits with column offsets): bitwise
    the oracle."""
    nx, ny = dims
    rng = np.random.default_rng(nx + ny)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    print(f"two-step shape={shape} dims={dims}", flush=True)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(7)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 7, O.COL_EXT_REG, 0.51, 0.65, nthreads=16))


def test_default_multi_step_on_ragged_width(gpu, orc):
    """A 1000 x 1000 periodic lattice (width no multiple of 64: neither the
    256-wide strips nor the tiles fit, so the default rule takes the
    three-step walk on strips of any width): 1 + 2 x 4 steps bitwise the
    oracle."""
    nx = ny = 1000
    rng = np.random.default_rng(7)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(9)
        assert lat.status() == (9, 0)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 9, O.COL_EXT_REG, 0.51, 0.65, nthreads=16))


@pytest.mark.parametrize("L,want", [(64, 3), (480, 3), (640, 7), (896, 47), (1152, 40)])
def test_default_rule_picks_bitwise(gpu, orc, L, want):
    """The shapes the default rule picks at small and mid sizes (32x8 tiles,
    64x4 tiles, the five-step walk below 1.5 workgroups per CU, the two-wavefront
    three-step walk between one and 1.5), for calls leaving every
    remainder: bitwise the strict oracle."""
    rng = np.random.default_rng(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    calls = (1, 4, 7, 5, 3, 2)
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        assert lat.step_shape()[0] == want
        for n in calls:
            lat.step(n)
        assert lat.status() == (sum(calls), 0)
        got = lat.download()
    assert np.array_equal(got, orc.run_steps(9, f0, st, sum(calls), O.COL_EXT_REG, 0.51, 0.65, nthreads=16))


@pytest.mark.parametrize("q,dims,seed", [(9, (300, 77, 1), 1), (9, (5, 3, 1), 2), (19, (37, 19, 11), 3),
                                         (19, (4, 3, 2), 4)])
def test_ragged_sizes_random_solids(gpu, orc, q, dims, seed):
    """Sizes that are no multiple of any block or tile, down to a few
    voxels, with random solids (bounce-back everywhere): bitwise the oracle,
    including zero-step calls and the state / post-collision downloads."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    m = (rng.uniform(size=n) > 0.3).astype(np.int32)
    m[0] = 1
    ini, st, nf = orc.stream_map(q, *dims, m)
    feq = orc.dist_eq(q, 0.02, -0.01, 0.01 if q == 19 else 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-3 * rng.uniform(-1, 1, (nf, q)))
    with gpu.Lattice(q, *dims, mask=m) as lat:
        lat.set_params(0.57, 0.68, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(0)
        assert lat.status() == (0, 0)
        assert np.array_equal(lat.download(), f0)
        lat.step(13)
        got = lat.download()
    want = orc.run_bc_steps(q, f0, *dims, ini, st, 13, O.COL_EXT_REG, 0.57, 0.68)[0]
    assert np.array_equal(got, want)


def test_lattice_errors_leave_the_handle_usable(gpu, orc):
    """Invalid calls return an error code with a message and change nothing."""
    L = 32
    f0 = gpu.init_double_shear(L)
    with gpu.Lattice(9, L, L, 1) as lat:
        with pytest.raises(RuntimeError, match="upload"):
            lat.step(1)
        with pytest.raises(ValueError):
            lat.upload(f0[:-1])
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        with pytest.raises(RuntimeError):
            lat.set_params(-1.0, 0.65, 1.0, O.COL_EXT_REG)
        with pytest.raises(RuntimeError):
            lat.set_params(0.6, 0.65, 1.0, 7)
        lat.upload(f0)
        with pytest.raises(RuntimeError):
            lat.download(gpu.POST_COLLISION)      # no step since the upload
        lat.step(3)
        ini, st, nf = orc.stream_map(9, L, L, 1)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 3, O.COL_EXT_REG, 0.6, 0.65))


def test_headline_default_kernel_full_size(gpu, orc):
    """The headline workload as the bench runs it (D2Q9 double shear,
    L = 2048, the default multi-step kernel: five steps per launch): 25
    steps bitwise the strict oracle; total mass conserved to rounding."""
    L, n = 2048, 25
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65, nthreads=16)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        shape, spl, desc = lat.step_shape()
        print(f"headline default: shape {shape}, {desc}", flush=True)
        assert spl == 5
        lat.step(n)
        got = lat.download()
    assert np.array_equal(got, want)
    assert abs(got.sum() - f0.sum()) <= 1e-12 * f0.sum()


@pytest.mark.parametrize("prio", ["0", "1", "2"])
def test_aligned_walk_issue_priority_bitwise(gpu, orc, monkeypatch, prio):
    """The lane-aligned walk with its issue priority by progress forced on
    and off (by default it is on where two or more workgroups share a CU):
    scheduling only, so bitwise the oracle either way."""
    nx, ny = 1000, 40
    rng = np.random.default_rng(5)
    ini, st, nf = orc.stream_map(9, nx, ny, 1)
    feq = orc.dist_eq(9, 0.03, -0.02, 0.0, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 9)))
    monkeypatch.setenv("ERLBM_FUSE", "42")
    monkeypatch.setenv("ERLBM_FUSE_PRIO", prio)
    with gpu.Lattice(9, nx, ny, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(10)
        assert np.array_equal(lat.download(), orc.run_steps(9, f0, st, 10, O.COL_EXT_REG, 0.51, 0.65, nthreads=16))


def test_default_shape_rule(gpu):
    """The default multi-step shape per lattice size (periodic FP64 D2Q9
    E-RLBM; gfx950 parts have 256 CUs): the five-step lane-aligned walk
    where it makes 1.5 workgroups per CU or more, else the three-step
    lane-aligned any-width walk where it loads the CUs no more unevenly
    than the 256-wide walk or makes three workgroups per CU, the 256-wide
    walk at L = 1024, the five-step walk from 0.85 workgroups per CU where
    256-wide strips do not divide the width (L = 896) and the two-wavefront
    three-step walk between one and 1.5 per CU (L = 1152), the 64x8
    two-step tiles at L = 512 .. 768, the 32x8
    ones below (fewer than two 64x8 workgroups per CU) and where 64x8 tiles
    do not fit (L = 480), and steps_per_launch / descriptions that match."""
    want = {64: (3, 2), 128: (3, 2), 256: (3, 2), 384: (3, 2), 448: (3, 2), 480: (3, 2), 512: (1, 2), 576: (1, 2),
            640: (7, 2), 704: (7, 2), 768: (1, 2), 896: (47, 5), 1000: (42, 3), 1024: (26, 3), 1152: (40, 3), 1280: (47, 5), 2048: (47, 5),
            3000: (47, 5), 4096: (47, 5)}
    words = {2: "two steps", 3: "three steps", 4: "four steps", 5: "five steps"}
    for L, (shape, spl) in want.items():
        with gpu.Lattice(9, L, L, 1) as lat:
            lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
            got = lat.step_shape()
            print(f"L={L}: {got}", flush=True)
            assert got[:2] == (shape, spl), L
            assert got[2].startswith(words[spl])
    # the moment-space arithmetic takes the five-step walk from 0.75
    # workgroups per CU (L = 768: 192 workgroups); MRT, which keeps the
    # reference's order, the shapes above
    for L, want_shape in ((512, (1, 2)), (768, (47, 5)), (1024, (47, 5)), (1280, (47, 5)), (2048, (47, 5)),
                          (4096, (47, 5))):
        with gpu.Lattice(9, L, L, 1) as lat:
            lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
            lat.set_arith(gpu.ARITH_MOMENT)
            got = lat.step_shape()
            assert got[:2] == want_shape, (L, got)
            lat.set_params(0.51, 0.65, 1.0, O.COL_MRT)
            assert lat.step_shape()[:2] == want.get(L, want_shape), L
    # shifted storage: FP32 takes the four-step lane-aligned walk (the
    # five-step one in the moment-space arithmetic), FP16 one step per launch
    for prec, want_shape in ((gpu.FP32_SHIFTED, (46, 4)), (gpu.FP16_SHIFTED, (0, 1))):
        with gpu.Lattice(9, 2048, 2048, 1, precision=prec) as lat:
            lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
            assert lat.step_shape()[:2] == want_shape, prec
    for L, want_shape in ((512, (0, 1)), (768, (47, 5)), (1024, (47, 5)), (1280, (47, 5)), (2048, (47, 5)),
                          (4096, (47, 5))):
        with gpu.Lattice(9, L, L, 1, precision=gpu.FP32_SHIFTED) as lat:
            lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
            lat.set_arith(gpu.ARITH_MOMENT)
            assert lat.step_shape()[:2] == want_shape, L
            lat.set_arith(gpu.ARITH_EXACT)
            assert lat.step_shape()[1] <= 4, L
    # a stop test since the upload: one step per launch
    with gpu.Lattice(9, 2048, 2048, 1) as lat:
        lat.set_params(0.51, 0.65, 1.0, O.COL_EXT_REG)
        lat.set_density_check(10.0)
        assert lat.step_shape()[:2] == (0, 1)


@pytest.mark.parametrize("shape", [11, 26, 42])
def test_multi_step_kernel_fma_tolerance_1000_steps(gpu, monkeypatch, shape):
    """The 1e-10 tolerance on the two- and three-step strip walks: max
    relative population error against the FMA-contracted oracle after 1000
    steps (D2Q9 double shear, L = 512; 256-wide strips, 42: lane-aligned
    strips of any width)."""
    import oracle
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    ofma = oracle.Oracle(fma=True)
    L = 512
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = ofma.stream_map(9, L, L, 1)
    want = ofma.run_steps(9, f0, st, 1000, O.COL_EXT_REG, tau, 0.65, nthreads=16)
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        lat.step(1000)
        got = lat.download()
    err = _rel(got, want)
    print(f"shape {shape} vs FMA oracle, L={L}, 1000 steps: max rel err {err:.2e}", flush=True)
    assert err <= TOL_FMA


@pytest.mark.parametrize("col", [O.COL_BGK, O.COL_REG, O.COL_EXT_REG])
@pytest.mark.parametrize("shape,dims", [(30, (64, 16, 40)), (31, (64, 8, 37)), (32, (96, 12, 33)), (33, (128, 16, 35)),
                                        (34, (128, 8, 33)), (35, (64, 32, 34))])
def test_d3q19_zwalk_bitwise(gpu, orc, monkeypatch, col, shape, dims):
    """D3Q19 two steps per launch (x-y tiles walked along z, segments of
    both walking directions, ragged segment lengths): bitwise the oracle
    for odd and even call lengths."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * ny + nz + col)
    ini, st, nf = orc.stream_map(19, nx, ny, nz)
    feq = orc.dist_eq(19, 0.02, -0.01, 0.015, 1.0)
    f0 = feq[None] * (1 + 1e-2 * rng.uniform(-1, 1, (nf, 19)))
    tau, tnh = 0.52, {O.COL_BGK: 0.52, O.COL_REG: 1.0, O.COL_EXT_REG: 0.65}[col]
    monkeypatch.setenv("ERLBM_FUSE", str(shape))
    print(f"d3q19 z-walk shape={shape} dims={dims} col={col}", flush=True)
    with gpu.Lattice(19, nx, ny, nz) as lat:
        lat.set_params(tau, tnh, 1.0, col)
        lat.upload(f0)
        done = 0
        for n in (1, 4, 3):
            lat.step(n)
            done += n
            assert lat.status() == (done, 0)
            assert np.array_equal(lat.download(), orc.run_steps(19, f0, st, done, col, tau, tnh, nthreads=16)), done


def test_d3q19_default_two_step_128(gpu, orc):
    """128^3 periodic D3Q19 E-RLBM (the default rule takes the z-walk there):
    1 + 2 x 2 steps bitwise the oracle; mass conserved to rounding."""
    n = 128
    rng = np.random.default_rng(5)
    ini, st, nf = orc.stream_map(19, n, n, n)
    f0 = orc.dist_eq(19, 0.02, -0.01, 0.015, 1.0)[None] * (1 + 1e-3 * rng.uniform(-1, 1, (nf, 19)))
    with gpu.Lattice(19, n, n, n) as lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lat.step(5)
        got = lat.download()
    assert np.array_equal(got, orc.run_steps(19, f0, st, 5, O.COL_EXT_REG, 0.6, 0.65, nthreads=16))
    assert abs(got.sum() - f0.sum()) <= 1e-12 * f0.sum()


def test_d3q19_zwalk_monitors_equal_one_step(gpu, monkeypatch):
    """After z-walk launches, the post-collision download and the bitwise
    kinetic-energy monitor equal those of the one-step path."""
    nx, ny, nz = 128, 16, 34
    f0 = gpu.dist_eq(19, 0.02, -0.01, 0.015, 1.0)[None] * (
        1 + 1e-2 * np.random.default_rng(11).uniform(-1, 1, (nx * ny * nz, 19)))
    lats = []
    for fuse in ("0", "33"):
        monkeypatch.setenv("ERLBM_FUSE", fuse)
        lat = gpu.Lattice(19, nx, ny, nz)
        lat.set_params(0.55, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        lats.append(lat)
    ref, fused = lats
    with ref, fused:
        for n in (1, 6, 5):
            ref.step(n)
            fused.step(n)
            assert fused.status() == ref.status()
            assert np.array_equal(fused.download(), ref.download())
            assert np.array_equal(fused.download(gpu.POST_COLLISION), ref.download(gpu.POST_COLLISION))
            assert fused.kinetic_energy(gpu.POST_COLLISION) == ref.kinetic_energy(gpu.POST_COLLISION)
            assert fused.kinetic_energy(gpu.STATE) == ref.kinetic_energy(gpu.STATE)


def test_two_step_remainder_bitwise(gpu, orc, monkeypatch):
    """Calls whose step count leaves two steps after the three-step launches
    run them as one two-step launch: at L = 2048 on the three-step
    lane-aligned walk, steps 4 (1 + 3), 5 (3 + 2), 2 (a lone two-step launch)
    and 1, bitwise the strict oracle (FP64) and the one-step path (FP32
    shifted storage).  (Remainders after four-step launches:
    test_gpu_walk4.py.)"""
    L, calls = 2048, (4, 5, 2, 1)
    tau = 0.5 + 3. * (1. / 32.) * L / 80000.
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    want = orc.run_steps(9, f0, st, sum(calls), O.COL_EXT_REG, tau, 0.65, nthreads=16)
    monkeypatch.setenv("ERLBM_FUSE", "42")
    with _periodic(gpu, 9, L, tau, 0.65, O.COL_EXT_REG, f0) as lat:
        assert lat.step_shape()[1] == 3
        for n in calls:
            lat.step(n)
        assert lat.status() == (sum(calls), 0)
        assert np.array_equal(lat.download(), want)
    monkeypatch.delenv("ERLBM_FUSE")
    out = []
    for fuse in (None, "0"):
        if fuse is not None:
            monkeypatch.setenv("ERLBM_FUSE", fuse)
        with gpu.Lattice(9, L, L, 1, precision=gpu.FP32_SHIFTED) as lat:
            lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
            lat.upload(f0)
            for n in calls:
                lat.step(n)
            out.append(lat.download())
    assert np.array_equal(out[0], out[1])
