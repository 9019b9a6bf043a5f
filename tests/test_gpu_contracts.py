"""Entry points and contracts of include/erlbm.h that the trajectory tests do
not reach on their own:

  * erlbm_density_monitor (density(), LBM_functions.cpp:5543-5552, over every
    fluid site): max |rho| and the violation flag against the oracle's
    per-site densities, including a state with one NaN population;
  * the Zou-He inlet's east branch (ZouHeD3Q19_vx_x with pos_x >= nx/2,
    Boundary_conditions.cpp:405-439) and devnull_rho_x on both sides, against
    the reference-built boundary fixtures;
  * "distinct handles may be driven from distinct host threads": two
    lattices stepped concurrently from two threads equal the same runs made
    one after the other.
"""
import threading

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _densities(orc, f):
    return np.array([orc.density(s) for s in f])


@pytest.mark.parametrize("case", ["d2q9_periodic", "d3q19_obstacle"])
def test_density_monitor_vs_oracle(gpu, orc, golden, case):
    rng = np.random.default_rng(11)
    if case == "d2q9_periodic":
        q, dims, mask = 9, (48, 40, 1), None
        ini, st, nf = orc.stream_map(9, 48, 40, 1)
        f0 = orc.dist_eq(9, 0.02, -0.01, 0.0, 1.0)[None] * (1 + 0.2 * rng.uniform(-1, 1, (nf, 9)))
        bc = {}
    else:
        q, dims, mask = 19, (22, 12, 1), gpu.obstacle_mask()
        ini, st, nf = orc.stream_map(19, 22, 12, 1, mask)
        f0 = orc.init_rest_d3q19(nf) * (1 + 0.2 * rng.uniform(-1, 1, (nf, 19)))
        bc = dict(inlet_x=1, u_in=0.05, outlet_x=20, rho_out=1.0)
    print(f"density monitor {case}", flush=True)
    with gpu.Lattice(q, *dims, mask=mask) as lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        if bc:
            lat.set_zouhe_inlet(bc["inlet_x"], bc["u_in"])
            lat.set_outlet_rescale(bc["outlet_x"], bc["rho_out"])
        lat.upload(f0)
        for step in range(2):
            for which in (gpu.STATE, gpu.POST_COLLISION):
                if which == gpu.POST_COLLISION and step == 0:
                    continue
                f = lat.download(which)
                rmax = float(np.max(np.abs(_densities(orc, f))))
                got, viol = lat.density_monitor(which, 0.0)
                assert got == rmax and not viol, (step, which)
                assert lat.density_monitor(which, np.nextafter(rmax, 0.0))[1]
                assert not lat.density_monitor(which, rmax)[1]
            lat.step(7)
    # one NaN population: flagged whatever the threshold; the maximum is
    # that of the other sites (the NaN site has no magnitude)
    bad = f0.copy()
    bad[nf // 3, 4] = np.nan
    with gpu.Lattice(q, *dims, mask=mask) as lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(bad)
        got, viol = lat.density_monitor(gpu.STATE, 0.0)
        d = _densities(orc, bad)
        assert viol and np.isnan(d[nf // 3])
        assert got == float(np.max(np.abs(d[np.isfinite(d)])))


@pytest.mark.parametrize("q,name,kind,pos,val", [(19, "zouhe_east", "inlet", 4, 0.05),
                                                 (19, "zouhe_west", "inlet", 1, 0.05),
                                                 (19, "devnull_east", "outlet", 4, 1.0),
                                                 (19, "devnull_west", "outlet", 1, 1.01),
                                                 (9, "zouhe", "inlet", 1, 0.05),
                                                 (9, "devnull_east", "outlet", 4, 1.0),
                                                 (9, "devnull_west", "outlet", 1, 1.01)])
def test_boundary_columns_vs_reference_fixture(gpu, orc, golden, q, name, kind, pos, val):
    """One step after an upload collides the boundary-corrected state in
    place, so the post-collision populations are the oracle's collision of
    the reference's boundary output (6 x 4 (x 2) channel with a solid voxel)."""
    g = golden("ref_functions.npz")
    nz = 2 if q == 19 else 1
    print(f"boundary column q={q} {name} at x={pos}", flush=True)
    with gpu.Lattice(q, 6, 4, nz, mask=g[f"bc_q{q}_mask"]) as lat:
        lat.set_params(0.6, 0.65, 1.0, O.COL_EXT_REG)
        if kind == "inlet":
            lat.set_zouhe_inlet(pos, val)
        else:
            lat.set_outlet_rescale(pos, val)
        lat.upload(g[f"bc_q{q}_in"])
        lat.step(1)
        got = lat.download(gpu.POST_COLLISION)
    want = orc.collide(g[f"bc_q{q}_{name}"], O.COL_EXT_REG, 0.6, 0.65)
    assert np.array_equal(got, want)


def test_zouhe_east_trajectory(gpu, orc):
    """A D3Q19 channel whose Zou-He column is in the east half (pos_x >=
    nx/2) and outlet in the west half, 200 steps, bitwise the oracle."""
    nx, ny = 22, 12
    m = np.array(gpu.obstacle_mask()).reshape(ny, nx)[:, ::-1].reshape(-1).copy()   # mirrored channel
    ini, st, nf = orc.stream_map(19, nx, ny, 1, m)
    f0 = orc.init_rest_d3q19(nf)
    want, n, v = orc.run_bc_steps(19, f0, nx, ny, 1, ini, st, 200, O.COL_EXT_REG, 0.56, 0.7, inlet_x=nx - 2,
                                  u_in=-0.04, outlet_x=1, rho_out=1.0)
    with gpu.Lattice(19, nx, ny, 1, mask=m) as lat:
        lat.set_params(0.56, 0.7, 1.0, O.COL_EXT_REG)
        lat.set_zouhe_inlet(nx - 2, -0.04)
        lat.set_outlet_rescale(1, 1.0)
        lat.upload(f0)
        lat.step(200)
        got = lat.download()
    assert n == 200 and np.array_equal(got, want)


def _run_dsl(gpu, L, n):
    tau = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(gpu.init_double_shear(L))
        lat.step(n)
        return lat.download(), lat.kinetic_energy(gpu.POST_COLLISION, ordered=True)


def _run_tg_map(gpu, n):
    with gpu.Lattice(19, 16, 16, 1) as lat:
        lat.set_params(0.5003, 1.0, 1.0, O.COL_REG)
        lat.upload(gpu.init_tg_map(16, 0.3))
        lat.step(n, energy_interval=100, energy_factor=2.0)
        return lat.download(), lat.status()


def test_two_handles_from_two_threads(gpu):
    want_a = _run_dsl(gpu, 256, 301)
    want_b = _run_tg_map(gpu, 700)
    got = {}

    def worker(key, fn):
        try:
            got[key] = fn()
        except Exception as e:   # surfaced below
            got[key] = e

    for rep in range(3):
        ta = threading.Thread(target=worker, args=("a", lambda: _run_dsl(gpu, 256, 301)))
        tb = threading.Thread(target=worker, args=("b", lambda: _run_tg_map(gpu, 700)))
        ta.start()
        tb.start()
        ta.join(60)
        tb.join(60)
        assert not ta.is_alive() and not tb.is_alive()
        assert not isinstance(got["a"], Exception), got["a"]
        assert not isinstance(got["b"], Exception), got["b"]
        assert np.array_equal(got["a"][0], want_a[0]) and got["a"][1] == want_a[1], rep
        assert np.array_equal(got["b"][0], want_b[0]) and got["b"][1] == want_b[1], rep
