"""Output fields and files of the reference drivers, from lattices in HBM.

Velocity and vorticity fields and the enstrophy are computed on the GPU
(erlbm_velocity_field / erlbm_vorticity_field / erlbm_enstrophy in
include/erlbm.h); the VTK files are written byte for byte as SimBoltz's
rec_velocity / rec_vorticity (SimBoltz_Functions/LBM_functions.cpp:4218-4346)
and the map drivers (src/main_stability-map_*.cpp) write them; the L2 errors
are those of tools/error_calculator.cpp:35-114.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from . import _lib as L
from ._lib import check, lib


def _dims(lat):
    return lat.nz, lat.ny, lat.nx


def velocity_field(lat, which: int = L.POST_COLLISION) -> np.ndarray:
    """u at every voxel (0 at solids), shape (nz, ny, nx, 3)."""
    out = np.empty(_dims(lat) + (3,), np.float64)
    check(lib().erlbm_velocity_field(lat._h, which, out.ctypes.data_as(C.c_void_p)), "erlbm_velocity_field")
    return out


def vorticity_field(lat, which: int = L.POST_COLLISION) -> np.ndarray:
    """Curl of u as rec_vorticity writes it, shape (nz, ny, nx, 3)."""
    out = np.empty(_dims(lat) + (3,), np.float64)
    check(lib().erlbm_vorticity_field(lat._h, which, out.ctypes.data_as(C.c_void_p)), "erlbm_vorticity_field")
    return out


def enstrophy(lat, which: int = L.POST_COLLISION) -> float:
    e = C.c_double()
    check(lib().erlbm_enstrophy(lat._h, which, C.byref(e)), "erlbm_enstrophy")
    return e.value


def _vec3(field: np.ndarray) -> Tuple[np.ndarray, int, int, int]:
    a = np.ascontiguousarray(field, np.float64)
    if a.ndim == 3:
        a = a[None]
    nz, ny, nx, three = a.shape
    if three != 3:
        raise ValueError("expected a (nz, ny, nx, 3) or (ny, nx, 3) vector field")
    return a, nx, ny, nz


def write_velocity_vtk(path: str, field: np.ndarray) -> None:
    a, nx, ny, nz = _vec3(field)
    check(lib().erlbm_write_velocity_vtk(path.encode(), nx, ny, nz, a.ctypes.data_as(C.c_void_p)),
          "erlbm_write_velocity_vtk")


def write_vorticity_vtk(path: str, field: np.ndarray) -> None:
    a, nx, ny, nz = _vec3(field)
    check(lib().erlbm_write_vorticity_vtk(path.encode(), nx, ny, nz, a.ctypes.data_as(C.c_void_p)),
          "erlbm_write_vorticity_vtk")


def velocity_vtk_name(t: float) -> str:
    """rec_velocity's file name: "vel_" + std::to_string(t) + ".vtk"."""
    return "vel_%f.vtk" % t


def read_velocity_vtk(path: str) -> np.ndarray:
    """The vectors of a rec_velocity file, shape (nz, ny, nx, 3)."""
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    check(lib().erlbm_read_velocity_vtk(path.encode(), None, 0, C.byref(nx), C.byref(ny), C.byref(nz)),
          "erlbm_read_velocity_vtk")
    out = np.empty((nz.value, ny.value, nx.value, 3), np.float64)
    check(lib().erlbm_read_velocity_vtk(path.encode(), out.ctypes.data_as(C.c_void_p), out.size, C.byref(nx),
                                        C.byref(ny), C.byref(nz)), "erlbm_read_velocity_vtk")
    return out


def vtk_round(a: np.ndarray) -> np.ndarray:
    """Values as a write-then-read of the VTK text returns them."""
    out = np.array(a, np.float64, copy=True, order="C")
    check(lib().erlbm_vtk_round(out.ctypes.data_as(C.c_void_p), out.size), "erlbm_vtk_round")
    return out


def l2_error(sim: np.ndarray, ref: np.ndarray) -> Tuple[float, float]:
    """(E_vector, E_magnitude) of a 2-D run against a reference run on an
    integer refinement of its grid (tools/error_calculator.cpp)."""
    s, nx, ny, nz = _vec3(sim)
    r, nxr, nyr, nzr = _vec3(ref)
    if nz != 1 or nzr != 1:
        raise ValueError("the error calculator compares 2-D fields")
    ev, em = C.c_double(), C.c_double()
    check(lib().erlbm_velocity_l2_error(s.ctypes.data_as(C.c_void_p), nx, ny, r.ctypes.data_as(C.c_void_p), nxr, nyr,
                                        C.byref(ev), C.byref(em)), "erlbm_velocity_l2_error")
    return ev.value, em.value


# ---- stability maps ----------------------------------------------------------------
def write_tg_map_vtk(path: str, stability: np.ndarray, collision: int, tau_nh: float) -> None:
    """stability: (n_nu, n_ma) ints, 1 = unstable (row = log10 nu axis)."""
    m = np.ascontiguousarray(stability, np.int32)
    check(lib().erlbm_write_tg_map_vtk(path.encode(), m.shape[1], m.shape[0], collision, tau_nh,
                                       m.ctypes.data_as(C.c_void_p)), "erlbm_write_tg_map_vtk")


def write_obstacle_map_vtk(path: str, stability: np.ndarray, ma: float = 0.1) -> None:
    m = np.ascontiguousarray(stability, np.int32)
    check(lib().erlbm_write_obstacle_map_vtk(path.encode(), m.shape[1], m.shape[0], ma,
                                             m.ctypes.data_as(C.c_void_p)), "erlbm_write_obstacle_map_vtk")


def read_map_vtk(path: str) -> np.ndarray:
    """A stability map file (the drivers' int format or the published files'
    tab-separated doubles) as an (n_j, n_i) int array, i fastest in the file."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    nx, ny = (int(v) for v in next(ln for ln in lines if ln.startswith("DIMENSIONS")).split()[1:3])
    k = next(n for n, ln in enumerate(lines) if ln.startswith("LOOKUP_TABLE"))
    vals = np.array(" ".join(lines[k + 1:]).split(), dtype=float)
    if vals.size < nx * ny:
        raise ValueError(f"{path}: {vals.size} values for a {nx} x {ny} map")
    return vals[: nx * ny].reshape(ny, nx).astype(np.int32)


def diff_maps(a: np.ndarray, b: np.ndarray) -> dict:
    """Point-by-point comparison of two maps of the same shape."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        raise ValueError(f"map shapes differ: {a.shape} vs {b.shape}")
    jj, ii = np.nonzero(a != b)
    return {"points": int(a.size), "agree": int(a.size - ii.size), "agreement": float(1 - ii.size / a.size),
            "unstable_a": int(a.sum()), "unstable_b": int(b.sum()),
            "differ": [(int(i), int(j)) for i, j in zip(ii, jj)]}
