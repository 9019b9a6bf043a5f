"""Output files and error measures (host side of erlbm_amd.fields), against
the reference's own output (tests/golden/ref_outputs.npz: rec_velocity,
rec_vorticity, the drivers' files and the error calculator, produced by the
reference compiled from its sources).  No GPU: the fields come from the
oracle's calcula or from the fixture files themselves."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "ref_outputs.npz")),
                                reason="output fixtures not generated")


def _bytes(g, key):
    return bytes(g[key])


def _vectors_from_text(text: bytes):
    lines = text.decode().split("\n")
    nx, ny, nz = (int(v) for v in lines[4].split()[1:4])
    vals = np.array(" ".join(lines[9:]).split(), dtype=float)
    return vals.reshape(nz, ny, nx, 3)


def _velocity_from_populations(orc, q, dims, mask, f):
    nx, ny, nz = dims
    out = np.zeros((nz * ny * nx, 3))
    fluid = np.nonzero(mask)[0]
    for p, v in enumerate(fluid):
        out[v] = orc.calcula(f[p])[:3]
    return out.reshape(nz, ny, nx, 3)


@pytest.mark.parametrize("q,dims", [(9, (12, 10, 1)), (19, (6, 5, 4))])
def test_velocity_file_bytes_equal_rec_velocity(orc, golden, tmp_path, q, dims):
    from erlbm_amd import fields
    g = golden("ref_outputs.npz")
    vel = _velocity_from_populations(orc, q, dims, g[f"q{q}_mask"], g[f"q{q}_f"])
    path = str(tmp_path / fields.velocity_vtk_name(0.25))
    fields.write_velocity_vtk(path, vel)
    assert os.path.basename(path) == "vel_0.250000.vtk"
    assert open(path, "rb").read() == _bytes(g, f"q{q}_vel")


@pytest.mark.parametrize("key", ["q9_vor", "q19_vor", "dsl32/vor_64.vtk"])
def test_vorticity_file_format(golden, tmp_path, key):
    """Values parsed from the reference's file, written again: same bytes
    (six significant digits survive the round trip)."""
    from erlbm_amd import fields
    g = golden("ref_outputs.npz")
    want = _bytes(g, key)
    path = str(tmp_path / "vor.vtk")
    fields.write_vorticity_vtk(path, _vectors_from_text(want))
    assert open(path, "rb").read() == want


def test_reader_and_round_trip(golden, tmp_path):
    from erlbm_amd import fields
    g = golden("ref_outputs.npz")
    p = tmp_path / "vel.vtk"
    p.write_bytes(_bytes(g, "dsl32/vel_0.125000.vtk"))
    v = fields.read_velocity_vtk(str(p))
    assert v.shape == (1, 32, 32, 3)
    assert np.array_equal(v, _vectors_from_text(_bytes(g, "dsl32/vel_0.125000.vtk")))
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 0.03, 1000), [0.0, -0.0, 1e-300, 123456.5, 1e-5, -7.25e-3]])
    r = fields.vtk_round(x)
    assert np.array_equal(r, np.array([float("%g" % a) for a in x]))
    assert np.array_equal(fields.vtk_round(r), r)          # idempotent


def test_error_calculator_matches_reference(golden, tmp_path):
    """tools/error_calculator on the L = 16 and L = 32 double-shear files:
    the same E_vector and E_magnitude to the ten digits it writes."""
    from erlbm_amd import fields
    g = golden("ref_outputs.npz")
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    a.write_bytes(_bytes(g, "dsl16/vel_0.125000.vtk"))
    b.write_bytes(_bytes(g, "dsl32/vel_0.125000.vtk"))
    ev, em = fields.l2_error(fields.read_velocity_vtk(str(a)), fields.read_velocity_vtk(str(b)))
    assert ("%d %.10g %.10g\n" % (16, ev, em)).encode() == _bytes(g, "error_calculator/erro_L2.dat")


def test_error_closed_forms():
    from erlbm_amd import fields
    rng = np.random.default_rng(2)
    ref = np.zeros((1, 32, 32, 3))
    ref[..., :2] = rng.normal(0, 0.05, (1, 32, 32, 2))
    sim = ref[:, ::2, ::2].copy()
    assert fields.l2_error(sim, ref) == (0.0, 0.0)
    ev, em = fields.l2_error(2 * sim, ref)                  # |2u - u| = |u|
    assert abs(ev - 1.0) < 1e-15 and abs(em - 1.0) < 1e-15
    for _ in range(5):
        s = sim + rng.normal(0, 0.02, sim.shape)
        ev, em = fields.l2_error(s, ref)
        assert em <= ev                                      # triangle inequality
    with pytest.raises(RuntimeError):
        fields.l2_error(ref[:, :30, :30], ref)               # not an integer refinement


@pytest.mark.parametrize("key,kind", [("map/tg.vtk", "tg"), ("map/obst.vtk", "obstacle")])
def test_map_files_bytes_equal_drivers(golden, tmp_path, key, kind):
    """The map drivers' files (4 x 3 Taylor-Green E-RLBM tau_nh 0.65, 3 x 4
    obstacle Ma 0.1): reading the map and writing it again gives the bytes."""
    from erlbm_amd import fields
    from erlbm_amd import _lib as L
    g = golden("ref_outputs.npz")
    want = _bytes(g, key)
    src = tmp_path / "src.vtk"
    src.write_bytes(want)
    m = fields.read_map_vtk(str(src))
    out = str(tmp_path / "out.vtk")
    if kind == "tg":
        fields.write_tg_map_vtk(out, m, L.COL_EXT_REG, 0.65)
    else:
        fields.write_obstacle_map_vtk(out, m, 0.1)
    assert open(out, "rb").read() == want
    d = fields.diff_maps(m, fields.read_map_vtk(out))
    assert d["agreement"] == 1.0 and d["differ"] == []


def test_published_map_reader():
    """The published maps (tab-separated doubles) read to the same 0/1 maps
    the bit-packed fixture holds."""
    from conftest import published_maps
    from erlbm_amd import fields
    maps = published_maps()
    m = maps["obstacle"]
    assert m.shape == (400, 400) and int(m.sum()) == 86407
    d = fields.diff_maps(m, m ^ (np.arange(m.size).reshape(m.shape) == 4001))
    assert d["agree"] == m.size - 1 and d["differ"] == [(1, 10)]
