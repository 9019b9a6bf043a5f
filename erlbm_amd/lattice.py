"""Host-side mirror of the SimBoltz lattice interface over the HIP C ABI.

The reference's drivers own ``GEOMETRY``/``LATTICE``/``PARAMETERS`` structs
(SimBoltz_Functions/Definitions.cpp:34-239) and call per-site functions in
their step loops.  ``Lattice`` keeps those concepts -- geometry mask, fluid
numbering, tau/tau_nh/rho_ini, collision choice, the two column boundary
conditions, the energy and density monitors -- but a step is one call into
``liberlbm.so`` (include/erlbm.h) that runs the whole sweep on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from ._lib import check, lib


class Lattice:
    """One lattice resident in HBM (FP64, or FP32/FP16 shifted deviations).

    q: 9 (D2Q9) or 19 (D3Q19).  mask[x + y nx + z nx ny] = 1 fluid / 0 solid
    (None: all fluid, periodic in every direction).  Populations cross in the
    reference's AoS fluid order, shape (n_fluid, q).
    """

    def __init__(self, q: int, nx: int, ny: int, nz: int = 1, mask: Optional[np.ndarray] = None,
                 stream: int = 0, precision: int = L.FP64):
        self.q, self.nx, self.ny, self.nz = q, nx, ny, nz
        h = C.c_void_p()
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32).reshape(-1)
        if m is not None and m.size != nx * ny * nz:
            raise ValueError("mask size does not match nx*ny*nz")
        mp = None if m is None else m.ctypes.data_as(C.c_void_p)
        self.precision = precision
        check(lib().erlbm_lattice_create(q, nx, ny, nz, mp, precision, C.c_void_p(stream or None),
                                         C.byref(h)), "erlbm_lattice_create")
        self._h = h
        self._mask = m
        n = C.c_int(0)
        check(lib().erlbm_lattice_fluid_count(h, C.byref(n)), "erlbm_lattice_fluid_count")
        self.n_fluid = n.value

    # -- lifecycle --------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().erlbm_lattice_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters / boundaries -----------------------------------------------
    def set_params(self, tau: float, tau_nh: float, rho0: float = 1.0, collision: int = L.COL_EXT_REG) -> None:
        check(lib().erlbm_set_params(self._h, tau, tau_nh, rho0, collision), "erlbm_set_params")

    def set_zouhe_inlet(self, pos_x: int, vx: float) -> None:
        check(lib().erlbm_set_zouhe_inlet(self._h, pos_x, vx), "erlbm_set_zouhe_inlet")

    def set_outlet_rescale(self, pos_x: int, rho: float) -> None:
        check(lib().erlbm_set_outlet_rescale(self._h, pos_x, rho), "erlbm_set_outlet_rescale")

    def set_density_check(self, limit: float) -> None:
        check(lib().erlbm_set_density_check(self._h, limit), "erlbm_set_density_check")

    def set_body_force(self, ax: float, ay: float, az: float = 0.0) -> None:
        """Body-force accelerations of the collision operators (the acc_x,
        acc_y, acc_z arguments of coll_* in Collision_operators.cpp)."""
        check(lib().erlbm_set_body_force(self._h, ax, ay, az), "erlbm_set_body_force")

    def set_arith(self, arith: int) -> None:
        """Collision arithmetic of the multi-step kernels: ARITH_EXACT (the
        reference's operation order, bitwise its strict build; the default)
        or ARITH_MOMENT (the same operator in moment space, within 1e-10 of
        the reference after 1000 steps, fewer FP64 instructions)."""
        check(lib().erlbm_set_arith(self._h, arith), "erlbm_set_arith")

    # -- transfer ----------------------------------------------------------------
    def upload(self, inif: np.ndarray) -> None:
        a = np.ascontiguousarray(inif, dtype=np.float64).reshape(-1)
        if a.size != self.n_fluid * self.q:
            raise ValueError(f"expected {self.n_fluid * self.q} populations, got {a.size}")
        check(lib().erlbm_upload_aos(self._h, a), "erlbm_upload_aos")

    def download(self, which: int = L.STATE) -> np.ndarray:
        out = np.empty(self.n_fluid * self.q, dtype=np.float64)
        check(lib().erlbm_download_aos(self._h, which, out), "erlbm_download_aos")
        return out.reshape(self.n_fluid, self.q)

    # -- stepping ----------------------------------------------------------------
    def step(self, n: int = 1, energy_interval: int = 0, energy_factor: float = 2.0) -> None:
        check(lib().erlbm_step(self._h, n, energy_interval, energy_factor), "erlbm_step")

    def step_timed(self, n: int) -> float:
        """n steps between two HIP events on the lattice's stream; milliseconds."""
        ms = C.c_float(0)
        check(lib().erlbm_step_timed(self._h, n, C.byref(ms)), "erlbm_step_timed")
        return ms.value

    def autotune(self, n_steps: int = 200):
        """Pick the fastest launch shape for this lattice (the run is not
        advanced); returns (description, microseconds per step)."""
        buf = C.create_string_buffer(128)
        us = C.c_float()
        check(lib().erlbm_lattice_autotune(self._h, n_steps, buf, 128, C.byref(us)), "erlbm_lattice_autotune")
        return buf.value.decode(), us.value

    def step_shape(self):
        """The multi-step kernel shape a plain step() call uses now: (shape,
        steps per launch, description); shape 0 = one step per launch."""
        f, spl = C.c_int(0), C.c_int(0)
        buf = C.create_string_buffer(128)
        check(lib().erlbm_step_shape(self._h, C.byref(f), C.byref(spl), buf, 128), "erlbm_step_shape")
        return f.value, spl.value, buf.value.decode()

    def status(self):
        s, r = C.c_longlong(0), C.c_int(0)
        check(lib().erlbm_status(self._h, C.byref(s), C.byref(r)), "erlbm_status")
        return s.value, r.value

    def synchronize(self) -> None:
        check(lib().erlbm_synchronize(self._h), "erlbm_synchronize")

    # -- monitors ----------------------------------------------------------------
    def kinetic_energy(self, which: int = L.STATE, ordered: bool = True) -> float:
        e = C.c_double(0)
        check(lib().erlbm_kinetic_energy(self._h, which, int(ordered), C.byref(e)), "erlbm_kinetic_energy")
        return e.value

    def density_monitor(self, which: int = L.STATE, thresh: float = 0.0):
        r, v = C.c_double(0), C.c_int(0)
        check(lib().erlbm_density_monitor(self._h, which, thresh, C.byref(r), C.byref(v)), "erlbm_density_monitor")
        return r.value, bool(v.value)


# ---- host-side driver helpers (C ABI, libm on the host) ------------------------
def autotune_candidates(q: int, nx: int, ny: int, nz: int = 1, precision: int = L.FP64, n_cu: int = 256) -> list:
    """The multi-step shapes (ERLBM_FUSE values) the autotuner would time for
    an all-fluid periodic lattice of this size, computed on the host."""
    cap = 256
    out = np.zeros(cap, np.int32)
    n = C.c_int(0)
    check(lib().erlbm_autotune_candidates(q, nx, ny, nz, precision, n_cu, out, cap, C.byref(n)),
          "erlbm_autotune_candidates")
    return [int(v) for v in out[: min(n.value, cap)]]


def dist_eq(q: int, vx: float, vy: float, vz: float, rho: float) -> np.ndarray:
    out = np.empty(q)
    check(lib().erlbm_dist_eq(q, vx, vy, vz, rho, out), "erlbm_dist_eq")
    return out


def init_taylor_green(q: int, L_: int, u0: float, tau: float, poisson: bool = True, neq: bool = True) -> np.ndarray:
    f = np.empty(L_ * L_ * q)
    check(lib().erlbm_init_taylor_green(q, L_, u0, tau, int(poisson), int(neq), f), "erlbm_init_taylor_green")
    return f.reshape(L_ * L_, q)


def init_double_shear(L_: int, rho: float = 1.0) -> np.ndarray:
    f = np.empty(L_ * L_ * 9)
    check(lib().erlbm_init_double_shear(L_, rho, f), "erlbm_init_double_shear")
    return f.reshape(L_ * L_, 9)


def init_tg_map(L_: int, u0: float, rho: float = 1.0) -> np.ndarray:
    f = np.empty(L_ * L_ * 19)
    check(lib().erlbm_init_tg_map(L_, u0, rho, f), "erlbm_init_tg_map")
    return f.reshape(L_ * L_, 19)


def obstacle_mask() -> np.ndarray:
    m = np.empty(22 * 12, np.int32)
    check(lib().erlbm_obstacle_mask(m), "erlbm_obstacle_mask")
    return m


def tg_map_point_params(i, j, n_ma=400, n_nu=400, L_=16, collision=L.COL_EXT_REG, tau_nh=0.65):
    u, t, tn, n = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    check(lib().erlbm_tg_map_point_params(i, j, n_ma, n_nu, L_, collision, tau_nh, C.byref(u), C.byref(t),
                                          C.byref(tn), C.byref(n)), "erlbm_tg_map_point_params")
    return u.value, t.value, tn.value, n.value


def obstacle_map_point_params(i, j, n_tau=400, n_nu=400, ma=0.1):
    u, t, tn, n = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    check(lib().erlbm_obstacle_map_point_params(i, j, n_tau, n_nu, ma, C.byref(u), C.byref(t), C.byref(tn),
                                                C.byref(n)), "erlbm_obstacle_map_point_params")
    return u.value, t.value, tn.value, n.value


def taylor_green_error(f_post: np.ndarray, L_: int, u0: float, visc: float, step: int) -> np.ndarray:
    """[E_vector, E_magnitude, (nu_eff - nu)/nu] of post-collision populations."""
    q = f_post.shape[-1]
    out = np.zeros(3)
    check(lib().erlbm_taylor_green_error(q, L_, u0, visc, step, np.ascontiguousarray(f_post, np.float64).reshape(-1),
                                         out), "erlbm_taylor_green_error")
    return out
