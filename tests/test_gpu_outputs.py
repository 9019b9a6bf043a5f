"""GPU output pipeline against the reference's own files and published tables.

* velocity / vorticity fields and enstrophy computed on the GPU, written
  with the VTK writers: byte-identical to rec_velocity / rec_vorticity and
  bitwise equal to calc_enstrophy (fixtures from the reference's build);
* the C++ drivers (erlbm_amd/bin, over the C ABI): the double-shear driver's
  whole output directory, the Taylor-Green driver's printed line and the map
  drivers' files byte-identical to the reference drivers' on the same
  arguments;
* the paper's convergence tables regenerated on the GPU: Fig. 3 (both
  initialisations) and Fig. 4 (against the RLBM L = 2048 reference run,
  65,537 steps) equal to data/*.csv to every printed digit; the kappa4 table
  equal to the reference's strict build and within one unit of the sixth
  digit of the published (FMA-built) table.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, read_csv

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "erlbm_amd", "bin")


def _fixtures(golden):
    if not os.path.exists(os.path.join(GOLDEN, "ref_outputs.npz")):
        pytest.skip("output fixtures not generated")
    return golden("ref_outputs.npz")


def _run(exe, args, cwd):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (python -m erlbm_amd.build)")
    print(f"running {exe} {' '.join(args)}", flush=True)
    p = subprocess.run([path, *args], cwd=cwd, capture_output=True, timeout=300)
    assert p.returncode == 0, p.stderr.decode()
    return p.stdout


@pytest.mark.parametrize("q,dims", [(9, (12, 10, 1)), (19, (6, 5, 4))])
def test_fields_and_enstrophy_equal_reference(gpu, golden, tmp_path, q, dims):
    from erlbm_amd import fields
    g = _fixtures(golden)
    with gpu.Lattice(q, *dims, mask=g[f"q{q}_mask"]) as lat:
        lat.set_params(0.6, 0.65, 1.0, gpu.COL_EXT_REG)
        lat.upload(g[f"q{q}_f"])
        vel = fields.velocity_field(lat, gpu.STATE)
        vor = fields.vorticity_field(lat, gpu.STATE)
        ens = fields.enstrophy(lat, gpu.STATE)
    fields.write_velocity_vtk(str(tmp_path / "v.vtk"), vel)
    fields.write_vorticity_vtk(str(tmp_path / "w.vtk"), vor)
    assert (tmp_path / "v.vtk").read_bytes() == bytes(g[f"q{q}_vel"])
    assert (tmp_path / "w.vtk").read_bytes() == bytes(g[f"q{q}_vor"])
    assert ens == float(g[f"q{q}_enstrophy"][0])


def test_double_shear_driver_files_identical(gpu, golden, tmp_path):
    """bin/double_shear_layer --L 16 / 32 --t-final 0.125 --dt-out 0.0625:
    k_energy_lbm.dat, vel_*.vtk and vor_*.vtk byte-identical to the
    reference driver's, and the error calculator's result on them."""
    from erlbm_amd import fields
    g = _fixtures(golden)
    names = [str(n) for n in g["dsl_files"]]
    for L in (16, 32):
        d = tmp_path / f"L{L}"
        d.mkdir()
        _run("double_shear_layer", ["--L", str(L), "--t-final", "0.125", "--dt-out", "0.0625"], str(d))
        want = sorted(n.split("/", 1)[1] for n in names if n.startswith(f"dsl{L}/"))
        assert sorted(os.listdir(d)) == want
        for fn in want:
            assert (d / fn).read_bytes() == bytes(g[f"dsl{L}/{fn}"]), (L, fn)
    ev, em = fields.l2_error(fields.read_velocity_vtk(str(tmp_path / "L16" / "vel_0.125000.vtk")),
                             fields.read_velocity_vtk(str(tmp_path / "L32" / "vel_0.125000.vtk")))
    assert ("%d %.10g %.10g\n" % (16, ev, em)).encode() == bytes(g["error_calculator/erro_L2.dat"])


def test_taylor_green_driver_line(gpu, golden, tmp_path):
    g = _fixtures(golden)
    out = _run("taylor_green", ["--L", "32", "--collision", "erlbm"], str(tmp_path))
    assert out == bytes(g["tg_default_line"])
    out = _run("taylor_green", ["--L", "32", "--tau", "0.6", "--u0", "0.001", "--collision", "reg", "--init", "eq"],
               str(tmp_path))
    assert out == bytes(g["tg_tau_line"])


def test_map_drivers_files_identical(gpu, golden, tmp_path):
    g = _fixtures(golden)
    _run("stability_map_taylor_green", ["--n-ma", "4", "--n-nu", "3", "--collision", "erlbm", "--tau-nh", "0.65",
                                        "--out", "tg.vtk"], str(tmp_path))
    _run("stability_map_obstacle", ["--n-tau", "3", "--n-nu", "4", "--out", "obst.vtk"], str(tmp_path))
    assert (tmp_path / "tg.vtk").read_bytes() == bytes(g["map/tg.vtk"])
    assert (tmp_path / "obst.vtk").read_bytes() == bytes(g["map/obst.vtk"])


@pytest.mark.parametrize("csv_name,init", [("fig3_taylor_green_Re50.csv", "neq"),
                                           ("fig3_taylor_green_Re50_feq_init.csv", "eq")])
def test_fig3_regenerated(gpu, csv_name, init):
    from erlbm_amd import convergence
    rows = convergence.fig3_rows(init=init)
    want = [",".join(r.values()) for r in read_csv(csv_name)]
    assert rows[1:] == want


def test_kappa4_table_regenerated(gpu):
    """All 96 runs of scripts/run_kappa4_check.sh: every printed digit equal
    to the reference's strict build (kappa4_check_L64_strict.csv, made with
    it); the published table came from the FMA-contracted build and differs
    from both in the sixth digit of 5 rows."""
    from erlbm_amd import convergence
    rows = convergence.kappa4_rows()
    strict = [",".join(r.values()) for r in read_csv("kappa4_check_L64_strict.csv")]
    published = [",".join(r.values()) for r in read_csv("kappa4_check_L64.csv")]
    assert len(rows) - 1 == len(strict) == len(published) == 96
    bad = [(a, b) for a, b in zip(rows[1:], strict) if a != b]
    assert not bad, bad[:5]
    same = sum(a == b for a, b in zip(rows[1:], published))
    print(f"kappa4: {same} / 96 rows equal to the published table", flush=True)
    assert same >= 91
    for a, b in zip(rows[1:], published):
        for x, y in zip(a.split(",")[7:], b.split(",")[7:]):
            assert abs(float(x) - float(y)) <= 2e-5 * abs(float(y)), (a, b)


def test_fig4_regenerated(gpu):
    """Every row of data/fig4_double_shear_Re80000.csv to every printed digit
    (the L = 1024 E-RLBM tau_nh = 0.75 row, which the reference's README
    attributes to an earlier code version, included)."""
    from erlbm_amd import convergence
    rows = convergence.fig4_rows()
    want = read_csv("fig4_double_shear_Re80000.csv")
    assert len(rows) - 1 == len(want) == 15
    for got, w in zip(rows[1:], want):
        m, L_, tau, tnh, ev, em = got.split(",")
        assert (m, int(L_), float(tau), float(tnh)) == (w["model"], int(w["L"]), float(w["tau"]), float(w["tau_nh"]))
        assert (ev, em) == (w["E_vector"], w["E_magnitude"]), got
