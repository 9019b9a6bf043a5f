"""The convergence studies of the reference, run on the GPU.

    python -m erlbm_amd.convergence fig3 [--init eq]   Taylor-Green, Re 50, tau 0.62
    python -m erlbm_amd.convergence kappa4             Taylor-Green L=64, U0=1e-3, tau sweep
    python -m erlbm_amd.convergence fig4               double shear layer, Re 80000, vs RLBM L=2048

Each case is the reference driver's run (src/main_taylor-green.cpp,
src/main_double-shear-layer.cpp) with the parameters its script passes
(scripts/run_fig3_taylor_green.sh, run_kappa4_check.sh,
run_fig4_double_shear_layer.sh), on a plain FP64 lattice; the error
measures are the drivers' own (Taylor-Green: main_taylor-green.cpp:178-213;
double shear: the VTK velocity files and tools/error_calculator.cpp).
Rows are printed as CSV with the scripts' columns.
"""
from __future__ import annotations

import argparse
import math
import sys
from typing import Dict, List, Optional

from . import _lib as L
from .fields import l2_error, velocity_field, vtk_round
from .lattice import Lattice, init_double_shear, init_taylor_green, taylor_green_error

MODELS = {"BGK": (L.COL_BGK, 1.0), "REG": (L.COL_REG, 1.0), "ERLBM065": (L.COL_EXT_REG, 0.65),
          "ERLBM075": (L.COL_EXT_REG, 0.75)}


def c_round(x: float) -> int:
    """C's round() (half away from zero) for x >= 0; x - floor(x) is exact."""
    r = math.floor(x)
    return int(r + 1 if x - r >= 0.5 else r)


def taylor_green_case(L_: int, collision: int, tau_nh_arg: float = 0.65, init: str = "neq",
                      density: str = "poisson", u0: Optional[float] = None, Re: float = 50.0) -> Dict:
    """One run of the Taylor-Green driver without --tau (tau from Re, run to
    t = T = L / U0): its printed line and the values."""
    u_0 = 2.0 / L_ if u0 is None else u0
    visc = u_0 * L_ / Re
    tau = 0.5 + 3.0 * visc
    tau_nh = 1.0 if collision == L.COL_REG else (tau if collision == L.COL_BGK else tau_nh_arg)
    nT = c_round(L_ / u_0)
    with Lattice(9, L_, L_, 1) as lat:
        lat.set_params(tau, tau_nh, 1.0, collision)
        lat.upload(init_taylor_green(9, L_, u_0, tau, poisson=density == "poisson", neq=init == "neq"))
        lat.step(nT + 1)
        ev, em, dnu = taylor_green_error(lat.download(L.POST_COLLISION), L_, u_0, visc, nT)
    line = "%d %.6g %.6g %.6g %d %.6g %.6g %.6g" % (L_, tau, tau_nh, u_0, nT, ev, em, dnu)
    return {"L": L_, "tau": tau, "tau_nh": tau_nh, "U0": u_0, "steps": nT, "E_vector": ev, "E_magnitude": em,
            "rel_visc_error": dnu, "line": line}


def fig3_rows(Ls=(32, 64, 128, 256), init: str = "neq") -> List[str]:
    """scripts/run_fig3_taylor_green.sh: model,L,tau,E_vector,E_magnitude."""
    rows = ["model,L,tau,E_vector,E_magnitude"]
    for L_ in Ls:
        for name in ("BGK", "REG", "ERLBM065"):
            col, tnh = MODELS[name]
            r = taylor_green_case(L_, col, tnh, init=init)
            f = r["line"].split()
            rows.append(f"{name},{f[0]},{f[1]},{f[5]},{f[6]}")
    return rows


def kappa4_rows(L_: int = 64, taus=(0.505, 0.51, 0.52, 0.535, 0.55, 0.575, 0.6, 0.625, 0.65, 0.7, 0.75, 0.8, 0.9,
                                    1.0, 1.1, 1.2)) -> List[str]:
    """scripts/run_kappa4_check.sh (tau given, U0 = 1e-3, 2 nu k^2 t = 1)."""
    rows = ["model,init,L,tau,tau_nh,U0,steps,E_vector,E_magnitude,rel_visc_error"]
    for tau in taus:
        for name, col, tnh in (("BGK", L.COL_BGK, tau), ("REG", L.COL_REG, 1.0), ("ERLBM065", L.COL_EXT_REG, 0.65)):
            for init in ("eq", "neq"):
                r = taylor_green_tau_case(L_, col, tnh, tau, 0.001, init)
                rows.append(f"{name},{init}," + r["line"].replace(" ", ","))
    return rows


def taylor_green_tau_case(L_: int, collision: int, tau_nh_arg: float, tau: float, u0: float,
                          init: str = "neq") -> Dict:
    """The driver with --tau (and --u0): run to 2 nu k^2 t = 1."""
    visc = (tau - 0.5) / 3.0
    tau_nh = 1.0 if collision == L.COL_REG else (tau if collision == L.COL_BGK else tau_nh_arg)
    k = 2.0 * math.pi / L_
    nT = c_round(1.0 / (2.0 * visc * k * k))
    with Lattice(9, L_, L_, 1) as lat:
        lat.set_params(tau, tau_nh, 1.0, collision)
        lat.upload(init_taylor_green(9, L_, u0, tau, poisson=True, neq=init == "neq"))
        lat.step(nT + 1)
        ev, em, dnu = taylor_green_error(lat.download(L.POST_COLLISION), L_, u0, visc, nT)
    line = "%d %.6g %.6g %.6g %d %.6g %.6g %.6g" % (L_, tau, tau_nh, u0, nT, ev, em, dnu)
    return {"steps": nT, "E_vector": ev, "E_magnitude": em, "rel_visc_error": dnu, "line": line}


def double_shear_tau(L_: int, Re: float = 80000.0) -> float:
    """The tau the Fig. 4 script passes: printf("%.12g", 0.5 + 3 U0 L / Re)."""
    return float("%.12g" % (0.5 + 3.0 * (1.0 / 32.0) * L_ / Re))


def double_shear_velocity(L_: int, collision: int, tau_nh: float, t: float = 1.0) -> "np.ndarray":  # noqa: F821
    """The velocity file the double-shear driver writes at t* = t
    (vel_1.000000.vtk for t = 1), as the error calculator reads it back."""
    U0 = 1.0 / 32.0
    tau = double_shear_tau(L_)
    tnh = 1.0 if collision == L.COL_REG else tau_nh
    steps_per_t = c_round(L_ / U0)
    n = c_round(t * steps_per_t)
    with Lattice(9, L_, L_, 1) as lat:
        lat.set_params(tau, tnh, 1.0, collision)
        lat.upload(init_double_shear(L_))
        lat.step(n + 1)
        vel = velocity_field(lat, L.POST_COLLISION)
    return vtk_round(vel)


def fig4_rows(Ls=(64, 128, 256, 512, 1024), ref_L: int = 2048) -> List[str]:
    """scripts/run_fig4_double_shear_layer.sh: RLBM/E-RLBM errors at t = T
    against the RLBM run at ref_L."""
    ref = double_shear_velocity(ref_L, L.COL_REG, 1.0)
    rows = ["model,L,tau,tau_nh,E_vector,E_magnitude"]
    for L_ in Ls:
        for name in ("REG", "ERLBM065", "ERLBM075"):
            col, tnh = MODELS[name]
            ev, em = l2_error(double_shear_velocity(L_, col, tnh), ref)
            rows.append(f"{name},{L_},{double_shear_tau(L_):.12g},{tnh:g},{ev:.6g},{em:.6g}")
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["fig3", "kappa4", "fig4"])
    ap.add_argument("--init", default="neq", choices=["eq", "neq"])
    a = ap.parse_args(argv)
    if a.study == "fig3":
        rows = fig3_rows(init=a.init)
    elif a.study == "kappa4":
        rows = kappa4_rows()
    else:
        rows = fig4_rows()
    print("\n".join(rows))
    return 0


if __name__ == "__main__":
    sys.exit(main())
