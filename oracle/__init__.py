"""CPU oracle for the E-RLBM collide-stream path -- TEST INFRASTRUCTURE.

Loads ``liboracle.so`` (the C restatement in ``erlbm_oracle.c``, built
strict with ``-ffp-contract=off``) and, where it was built, the reference's
own library compiled from source into ``oracle/_ref``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package, as the checker; the product path (``erlbm_amd``) never does.

Parity pins of every function here are in ``tests/test_oracle_pins.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

COL_BGK, COL_MRT, COL_REG, COL_EXT_REG = 1, 2, 3, 4

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d, _i = C.c_double, C.c_int


def build(ref: bool = True) -> None:
    """Compile liboracle(.so, _fma.so); also oracle/_ref when the reference
    checkout exists (this container only)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/SimBoltz_Functions"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class _Lib:
    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run oracle.build())")
        self.lib = C.CDLL(path)


class Oracle(_Lib):
    """The C restatement (strict by default; ``fma=True`` for the contracted build)."""

    def __init__(self, fma: bool = False):
        super().__init__(os.path.join(HERE, "liboracle_fma.so" if fma else "liboracle.so"))
        L = self.lib
        for s, q in (("d2q9", 9), ("d3q19", 19)):
            _sig(L, f"orc_lattice_{s}", None, [_dp, _dp, _dp, _ip])
            _sig(L, f"orc_calcula_{s}", None, [_dp, _dp])
            _sig(L, f"orc_dist_eq_{s}", None, [_dp, _d, _d, _d, _d])
            _sig(L, f"orc_density_{s}", _d, [_dp])
            _sig(L, f"orc_collide_{s}", None, [_dp, _i, _d, _d, _d, _i, _d, _d])
            _sig(L, f"orc_number_fluid_{s}", _i, [_ip, _i, _ip])
            _sig(L, f"orc_stream_map_{s}", None, [_i, _i, _i, _ip, _ip])
            _sig(L, f"orc_propagate_{s}", None, [_dp, _dp, _ip, _i])
            _sig(L, f"orc_kinetic_energy_{s}", _d, [_dp, _i])
            _sig(L, f"orc_devnull_rho_x_{s}", None, [_i, _d, _i, _i, _i, _ip, _dp])
            _sig(L, f"orc_collide_stream_{s}", _i, [_dp, _dp, _ip, _i, _i, _d, _d, _d, _i])
            _sig(L, f"orc_collide_stream_acc_{s}", _i, [_dp, _dp, _ip, _i, _i, _d, _d, _d, _d, _d, _d, _i])
            _sig(L, f"orc_run_steps_{s}", None, [_dp, _i, _ip, _i, _i, _d, _d, _d, _d, _d, _i])
        _sig(L, "orc_zouhe_d3q19_vx_x", None, [_i, _d, _i, _i, _i, _ip, _dp])
        _sig(L, "orc_zouhe_d2q9_vx_x", None, [_i, _d, _i, _i, _ip, _dp])
        _sig(L, "orc_init_taylor_green_d2q9", None, [_i, _d, _d, _i, _i, _dp])
        _sig(L, "orc_init_double_shear_d2q9", None, [_i, _d, _dp])
        _sig(L, "orc_init_tg_map_d3q19", None, [_i, _d, _d, _dp])
        _sig(L, "orc_init_rest_d3q19", None, [_i, _d, _dp])
        _sig(L, "orc_obstacle_mask", None, [_ip])
        _sig(L, "orc_tg_map_params", None, [_i, _i, _i, _i, _i, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)])
        _sig(L, "orc_obstacle_map_params", None, [_i, _i, _i, _i, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)])
        _sig(L, "orc_run_tg_map_point", _i, [_i, _i, _d, _d, _d, C.POINTER(_i), C.POINTER(_d)])
        _sig(L, "orc_run_obstacle_map_point", _i, [_d, _d, _d, C.POINTER(_i)])
        _sig(L, "orc_run_map", None, [_i, _i, _i, _d, _d, _i, _i, _ip, _ip, _i, _ip, _ip, _i])
        _sig(L, "orc_tg_error_d2q9", None, [_dp, _i, _d, _d, _i, _dp])
        _sig(L, "orc_run_taylor_green_d2q9", None, [_i, _i, _d, _d, _d, _d, _i, _i, _i, _dp])
        _sig(L, "orc_run_bc_steps_d3q19", _i, [_dp, _i, _i, _i, _ip, _ip, _i, _i, _i, _d, _d, _i, _d, _i, _d, _d,
                                               _d, _d, _d, C.POINTER(_i)])
        _sig(L, "orc_run_bc_steps_d2q9", _i, [_dp, _i, _i, _ip, _ip, _i, _i, _i, _d, _d, _i, _d, _i, _d, _d,
                                              _d, _d, _d, C.POINTER(_i)])
        _sig(L, "orc_max_threads", _i, [])

    # ---- lattice constants -------------------------------------------------
    def lattice(self, q: int):
        s = _s(q)
        d = 2 if q == 9 else 3
        c = np.zeros(q * d)
        w = np.zeros(q)
        Q = np.zeros(q * d * d)
        opp = np.zeros(q, np.int32)
        getattr(self.lib, f"orc_lattice_{s}")(c, w, Q, opp)
        return c.reshape(q, d), w, Q, opp

    # ---- per-site functions ------------------------------------------------
    def collide(self, f: np.ndarray, col: int, tau: float, tau_nh: float, acc=(0.0, 0.0, 0.0)) -> np.ndarray:
        q = f.shape[-1]
        g = np.ascontiguousarray(f, dtype=np.float64).copy()
        getattr(self.lib, f"orc_collide_{_s(q)}")(g.reshape(-1), g.size // q, acc[0], acc[1], acc[2], col, tau, tau_nh)
        return g

    def calcula(self, f: np.ndarray) -> np.ndarray:
        out = np.zeros(4)
        getattr(self.lib, f"orc_calcula_{_s(f.size)}")(np.ascontiguousarray(f, np.float64), out)
        return out

    def dist_eq(self, q: int, vx, vy, vz, rho) -> np.ndarray:
        out = np.zeros(q)
        getattr(self.lib, f"orc_dist_eq_{_s(q)}")(out, vx, vy, vz, rho)
        return out

    def density(self, f: np.ndarray) -> float:
        return getattr(self.lib, f"orc_density_{_s(f.size)}")(np.ascontiguousarray(f, np.float64))

    # ---- geometry / streaming ---------------------------------------------
    def stream_map(self, q: int, nx: int, ny: int, nz: int, mask=None):
        n = nx * ny * nz
        m = np.ones(n, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32).reshape(-1)
        ini = np.zeros(n, np.int32)
        nf = getattr(self.lib, f"orc_number_fluid_{_s(q)}")(m, n, ini)
        st = np.zeros(nf * q, np.int32)
        getattr(self.lib, f"orc_stream_map_{_s(q)}")(nx, ny, nz, ini, st)
        return ini, st, nf

    def propagate(self, q: int, inif: np.ndarray, stream: np.ndarray) -> np.ndarray:
        out = np.zeros_like(inif)
        getattr(self.lib, f"orc_propagate_{_s(q)}")(inif.reshape(-1), out.reshape(-1), stream, inif.size // q)
        return out

    def kinetic_energy(self, q: int, inif: np.ndarray) -> float:
        return getattr(self.lib, f"orc_kinetic_energy_{_s(q)}")(np.ascontiguousarray(inif).reshape(-1), inif.size // q)

    # ---- boundaries --------------------------------------------------------
    def devnull_rho_x(self, q, pos_x, rho, nx, ny, nz, ini, inif):
        getattr(self.lib, f"orc_devnull_rho_x_{_s(q)}")(pos_x, rho, nx, ny, nz, ini, inif.reshape(-1))

    def zouhe_d3q19(self, pos_x, vx, nx, ny, nz, ini, inif):
        self.lib.orc_zouhe_d3q19_vx_x(pos_x, vx, nx, ny, nz, ini, inif.reshape(-1))

    def zouhe_d2q9(self, pos_x, vx, nx, ny, ini, inif):
        self.lib.orc_zouhe_d2q9_vx_x(pos_x, vx, nx, ny, ini, inif.reshape(-1))

    # ---- step loops ----------------------------------------------------------
    def collide_stream(self, q, a, b, stream, col, tau, tau_nh, dens_limit=0.0, nthreads=1) -> int:
        return getattr(self.lib, f"orc_collide_stream_{_s(q)}")(a.reshape(-1), b.reshape(-1), stream, a.size // q,
                                                               col, tau, tau_nh, dens_limit, nthreads)

    def run_steps(self, q, inif, stream, n_steps, col, tau, tau_nh, nthreads=1, acc=(0.0, 0.0, 0.0)) -> np.ndarray:
        """n_steps collide-stream sweeps; acc = the collisions' body-force
        accelerations (ax, ay, az)."""
        a = np.ascontiguousarray(inif, np.float64).copy()
        getattr(self.lib, f"orc_run_steps_{_s(q)}")(a.reshape(-1), a.size // q, stream, n_steps, col, tau, tau_nh,
                                                    float(acc[0]), float(acc[1]), float(acc[2]), nthreads)
        return a

    def run_bc_steps(self, q, inif, nx, ny, nz, ini, stream, n_steps, col, tau, tau_nh, inlet_x=-1, u_in=0.0,
                     outlet_x=-1, rho_out=1.0, dens_limit=0.0, acc=(0.0, 0.0, 0.0)):
        a = np.ascontiguousarray(inif, np.float64).copy()
        viol = C.c_int(0)
        ax, ay, az = (float(v) for v in acc)
        if q == 19:
            n = self.lib.orc_run_bc_steps_d3q19(a.reshape(-1), nx, ny, nz, ini, stream, a.size // q, n_steps, col,
                                                tau, tau_nh, inlet_x, u_in, outlet_x, rho_out, dens_limit,
                                                ax, ay, az, C.byref(viol))
        else:
            n = self.lib.orc_run_bc_steps_d2q9(a.reshape(-1), nx, ny, ini, stream, a.size // q, n_steps, col, tau,
                                               tau_nh, inlet_x, u_in, outlet_x, rho_out, dens_limit, ax, ay, az,
                                               C.byref(viol))
        return a, n, viol.value

    # ---- driver inits and runs ---------------------------------------------
    def init_taylor_green_d2q9(self, L, u0, tau, poisson=True, neq=True) -> np.ndarray:
        f = np.zeros((L * L, 9))
        self.lib.orc_init_taylor_green_d2q9(L, u0, tau, int(poisson), int(neq), f.reshape(-1))
        return f

    def init_double_shear_d2q9(self, L, rho=1.0) -> np.ndarray:
        f = np.zeros((L * L, 9))
        self.lib.orc_init_double_shear_d2q9(L, rho, f.reshape(-1))
        return f

    def init_tg_map_d3q19(self, L, u0, rho=1.0) -> np.ndarray:
        f = np.zeros((L * L, 19))
        self.lib.orc_init_tg_map_d3q19(L, u0, rho, f.reshape(-1))
        return f

    def init_rest_d3q19(self, n_fluid, rho=1.0) -> np.ndarray:
        f = np.zeros((n_fluid, 19))
        self.lib.orc_init_rest_d3q19(n_fluid, rho, f.reshape(-1))
        return f

    def obstacle_mask(self) -> np.ndarray:
        m = np.zeros(22 * 12, np.int32)
        self.lib.orc_obstacle_mask(m)
        return m

    def tg_map_params(self, i, j, n_ma=400, n_nu=400, col=COL_EXT_REG, tau_nh=0.65):
        a, b, c = _d(), _d(), _d()
        self.lib.orc_tg_map_params(i, j, n_ma, n_nu, int(col == COL_REG), tau_nh, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def obstacle_map_params(self, i, j, n_tau=400, n_nu=400, ma=0.1):
        a, b, c = _d(), _d(), _d()
        self.lib.orc_obstacle_map_params(i, j, n_tau, n_nu, ma, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def run_tg_map_point(self, L, col, u0, tau, tau_nh):
        st, e = _i(), _d()
        r = self.lib.orc_run_tg_map_point(L, col, u0, tau, tau_nh, C.byref(st), C.byref(e))
        return r, st.value, e.value

    def run_obstacle_map_point(self, u_imp, tau, tau_nh):
        st = _i()
        r = self.lib.orc_run_obstacle_map_point(u_imp, tau, tau_nh, C.byref(st))
        return r, st.value

    def run_map(self, kind, ii, jj, n_i=400, n_j=400, col=COL_EXT_REG, tau_nh=0.65, ma=0.1, L=16, nthreads=0):
        ii = np.ascontiguousarray(ii, np.int32)
        jj = np.ascontiguousarray(jj, np.int32)
        out = np.zeros(ii.size, np.int32)
        steps = np.zeros(ii.size, np.int32)
        nt = nthreads if nthreads > 0 else self.lib.orc_max_threads()
        self.lib.orc_run_map(kind, L, col, tau_nh, ma, n_i, n_j, ii, jj, ii.size, out, steps, nt)
        return out, steps

    def tg_error_d2q9(self, inif, L, u0, visc, step):
        out = np.zeros(3)
        self.lib.orc_tg_error_d2q9(np.ascontiguousarray(inif).reshape(-1), L, u0, visc, step, out)
        return out

    def run_taylor_green_d2q9(self, L=64, col=COL_EXT_REG, tau_nh=0.65, Re=50.0, u0=None, tau=-1.0, neq=True,
                              poisson=True, nthreads=0):
        out = np.zeros(5)
        u0 = 2.0 / L if u0 is None else u0
        nt = nthreads if nthreads > 0 else self.lib.orc_max_threads()
        self.lib.orc_run_taylor_green_d2q9(L, col, tau_nh, Re, u0, tau, int(neq), int(poisson), nt, out)
        return out

    def max_threads(self) -> int:
        return self.lib.orc_max_threads()


class Reference(_Lib):
    """The reference's own library, compiled from its sources into oracle/_ref."""

    def __init__(self, q: int, fma: bool = False):
        self.q = q
        self.s = _s(q)
        super().__init__(os.path.join(REF_DIR, f"libref_{self.s}_{'fma' if fma else 'strict'}.so"))
        L, s = self.lib, self.s
        _sig(L, f"ref_lattice_{s}", None, [_dp, _dp, _dp])
        _sig(L, f"ref_stream_map_{s}", _i, [_i, _i, _i, _ip, _ip, _ip])
        _sig(L, f"ref_collide_{s}", None, [_dp, _i, _d, _d, _d, _i, _d, _d])
        _sig(L, f"ref_calcula_{s}", None, [_dp, _dp])
        _sig(L, f"ref_dist_eq_{s}", None, [_dp, _d, _d, _d, _d])
        _sig(L, f"ref_density_{s}", _d, [_dp])
        _sig(L, f"ref_kinetic_energy_{s}", _d, [_i, _i, _i, _ip, _dp])
        _sig(L, f"ref_boundary_{s}", None, [_i, _i, _i, _i, _ip, _dp, _i, _d])
        _sig(L, f"ref_run_steps_{s}", _i, [_i, _i, _i, _ip, _dp, _i, _i, _d, _d, _i, _d, _i, _d, _d, _d, _d, _d,
                                           C.POINTER(_i)])
        _sig(L, f"ref_run_steps_omp_{s}", _d, [_i, _i, _i, _dp, _i, _i, _d, _d, _i])
        _sig(L, f"ref_rec_velocity_{s}", None, [_i, _i, _i, _ip, _dp, _d])
        _sig(L, f"ref_rec_vorticity_{s}", None, [_i, _i, _i, _ip, _dp, C.c_uint])
        _sig(L, f"ref_enstrophy_{s}", _d, [_i, _i, _i, _ip, _dp])
        _sig(L, f"ref_driver_main_{s}", _i, [_i, _i, C.POINTER(C.c_char_p)])
        if q == 9:
            _sig(L, "ref_init_taylor_green_d2q9", None, [_i, _d, _d, _i, _i, _dp])
            _sig(L, "ref_init_double_shear_d2q9", None, [_i, _dp])
        else:
            _sig(L, "ref_tg_map_points", None, [_i, _d, _i, _i, _i, _ip, _ip, _i, _ip])
            _sig(L, "ref_obstacle_map_points", None, [_d, _i, _i, _ip, _ip, _i, _ip])
            _sig(L, "ref_obstacle_geometry", _i, [_ip])

    @staticmethod
    def available(q: int = 9, fma: bool = False) -> bool:
        return os.path.exists(os.path.join(REF_DIR, f"libref_{_s(q)}_{'fma' if fma else 'strict'}.so"))

    def lattice(self):
        d = 2 if self.q == 9 else 3
        c, w, Q = np.zeros(self.q * d), np.zeros(self.q), np.zeros(self.q * d * d)
        getattr(self.lib, f"ref_lattice_{self.s}")(c, w, Q)
        return c.reshape(self.q, d), w, Q

    def stream_map(self, nx, ny, nz, mask=None):
        n = nx * ny * nz
        m = np.ones(n, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32).reshape(-1)
        ini = np.zeros(n, np.int32)
        st = np.zeros(n * self.q, np.int32)
        nf = getattr(self.lib, f"ref_stream_map_{self.s}")(nx, ny, nz, m, ini, st)
        return ini, st[: nf * self.q].copy(), nf

    def collide(self, f, col, tau, tau_nh, acc=(0.0, 0.0, 0.0)):
        g = np.ascontiguousarray(f, np.float64).copy()
        getattr(self.lib, f"ref_collide_{self.s}")(g.reshape(-1), g.size // self.q, acc[0], acc[1], acc[2], col, tau,
                                                   tau_nh)
        return g

    def calcula(self, f):
        out = np.zeros(4)
        getattr(self.lib, f"ref_calcula_{self.s}")(np.ascontiguousarray(f, np.float64).copy(), out)
        return out

    def dist_eq(self, vx, vy, vz, rho):
        out = np.zeros(self.q)
        getattr(self.lib, f"ref_dist_eq_{self.s}")(out, vx, vy, vz, rho)
        return out

    def density(self, f):
        return getattr(self.lib, f"ref_density_{self.s}")(np.ascontiguousarray(f, np.float64).copy())

    def kinetic_energy(self, nx, ny, nz, mask, inif):
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        return getattr(self.lib, f"ref_kinetic_energy_{self.s}")(nx, ny, nz, m, np.ascontiguousarray(inif).reshape(-1))

    def boundary(self, which, nx, ny, nz, mask, inif, pos_x, value):
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        g = np.ascontiguousarray(inif, np.float64).copy()
        getattr(self.lib, f"ref_boundary_{self.s}")(which, nx, ny, nz, m, g.reshape(-1), pos_x, value)
        return g

    def run_steps(self, nx, ny, nz, mask, inif, n_steps, col, tau, tau_nh, inlet_x=-1, u_in=0.0, outlet_x=-1,
                  rho_out=1.0, dens_limit=0.0, acc=(0.0, 0.0, 0.0)):
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        g = np.ascontiguousarray(inif, np.float64).copy()
        viol = C.c_int(0)
        ax, ay, az = (float(v) for v in acc)
        n = getattr(self.lib, f"ref_run_steps_{self.s}")(nx, ny, nz, m, g.reshape(-1), n_steps, col, tau, tau_nh,
                                                         inlet_x, u_in, outlet_x, rho_out, dens_limit, ax, ay, az,
                                                         C.byref(viol))
        return g, n, viol.value

    def run_steps_omp(self, nx, ny, nz, inif, n_steps, col, tau, tau_nh, nthreads=0) -> float:
        """The double-shear driver's OpenMP step loop, in place on a periodic
        lattice; returns wall seconds."""
        a = inif if inif.flags["C_CONTIGUOUS"] and inif.dtype == np.float64 else np.ascontiguousarray(inif, np.float64)
        t = getattr(self.lib, f"ref_run_steps_omp_{self.s}")(nx, ny, nz, a.reshape(-1), n_steps, col, tau, tau_nh,
                                                             nthreads)
        if a is not inif:
            inif[...] = a
        return t

    def init_taylor_green(self, L, u0, tau, poisson=True, neq=True):
        f = np.zeros((L * L, 9))
        self.lib.ref_init_taylor_green_d2q9(L, u0, tau, int(poisson), int(neq), f.reshape(-1))
        return f

    def init_double_shear(self, L):
        f = np.zeros((L * L, 9))
        self.lib.ref_init_double_shear_d2q9(L, f.reshape(-1))
        return f

    def tg_map_points(self, col, tau_nh, ii, jj, n_ma=400, n_nu=400, L=16):
        ii = np.ascontiguousarray(ii, np.int32)
        jj = np.ascontiguousarray(jj, np.int32)
        out = np.zeros(ii.size, np.int32)
        self.lib.ref_tg_map_points(col, tau_nh, L, n_ma, n_nu, ii, jj, ii.size, out)
        return out

    def obstacle_map_points(self, ii, jj, n_tau=400, n_nu=400, ma=0.1):
        ii = np.ascontiguousarray(ii, np.int32)
        jj = np.ascontiguousarray(jj, np.int32)
        out = np.zeros(ii.size, np.int32)
        self.lib.ref_obstacle_map_points(ma, n_tau, n_nu, ii, jj, ii.size, out)
        return out

    def rec_velocity(self, nx, ny, nz, mask, inif, tempo, cwd):
        """rec_velocity into directory cwd; returns the file's bytes."""
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        g = np.ascontiguousarray(inif, np.float64).copy()
        with _InDir(cwd):
            getattr(self.lib, f"ref_rec_velocity_{self.s}")(nx, ny, nz, m, g.reshape(-1), tempo)
        with open(os.path.join(cwd, "vel_%f.vtk" % tempo), "rb") as fh:
            return fh.read()

    def rec_vorticity(self, nx, ny, nz, mask, inif, passo, cwd):
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        g = np.ascontiguousarray(inif, np.float64).copy()
        with _InDir(cwd):
            getattr(self.lib, f"ref_rec_vorticity_{self.s}")(nx, ny, nz, m, g.reshape(-1), passo)
        with open(os.path.join(cwd, f"vor_{passo}.vtk"), "rb") as fh:
            return fh.read()

    def enstrophy(self, nx, ny, nz, mask, inif):
        m = np.ones(nx * ny * nz, np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
        return getattr(self.lib, f"ref_enstrophy_{self.s}")(nx, ny, nz, m,
                                                             np.ascontiguousarray(inif, np.float64).reshape(-1))

    def driver_main(self, which, args, cwd):
        """A reference driver's main() run in directory cwd; returns (status,
        what it printed to stdout).  D2Q9: 0 taylor-green, 1 double shear
        layer, 2 error calculator; D3Q19: 0 Taylor-Green map, 1 obstacle map."""
        argv = (C.c_char_p * (len(args) + 1))(*[a.encode() for a in ["driver", *args]])
        with _InDir(cwd), _Stdout(os.path.join(cwd, "_stdout.txt")):
            rc = getattr(self.lib, f"ref_driver_main_{self.s}")(which, len(args) + 1, argv)
        with open(os.path.join(cwd, "_stdout.txt"), "rb") as fh:
            out = fh.read()
        os.remove(os.path.join(cwd, "_stdout.txt"))
        return rc, out

    def obstacle_geometry(self):
        ini = np.zeros(22 * 12, np.int32)
        n = self.lib.ref_obstacle_geometry(ini)
        return ini, n


class _InDir:
    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.d)

    def __exit__(self, *exc):
        os.chdir(self.old)


class _Stdout:
    """Redirect file descriptor 1 (C printf / std::cout) to a file."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.fd = os.open(self.path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        os.dup2(self.fd, 1)

    def __exit__(self, *exc):
        _libc_fflush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.fd)


def _libc_fflush():
    try:
        C.CDLL(None).fflush(None)
    except Exception:
        pass


def _s(q: int) -> str:
    if q == 9:
        return "d2q9"
    if q == 19:
        return "d3q19"
    raise ValueError(f"unsupported velocity set q={q}")
