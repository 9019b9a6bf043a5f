"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, the ctypes table binds them all, and the host-side
driver helpers (initial conditions, map axes, geometry) equal the oracle's
bit for bit."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

LIB = os.path.join(ROOT, "erlbm_amd", "liberlbm.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(erlbm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_exports_every_declared_symbol():
    import erlbm_amd
    erlbm_amd.lib()  # loads, binds every signature
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = decl - exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"
    from erlbm_amd._lib import SIGNATURES
    unbound = decl - set(SIGNATURES)
    assert not unbound, f"declared but not bound in erlbm_amd._lib: {sorted(unbound)}"


def test_device_count_without_gpu_is_not_an_error():
    import erlbm_amd
    assert erlbm_amd.device_count() >= 0


def test_null_handles_are_argument_errors():
    """Calls on a null handle fail with ERLBM_ERR_ARG before any HIP call
    (so this runs without a GPU)."""
    import ctypes as C
    import erlbm_amd
    from erlbm_amd._lib import ERLBM_ERR_ARG
    lib = erlbm_amd.lib()
    f, spl = C.c_int(-1), C.c_int(-1)
    assert lib.erlbm_step_shape(None, C.byref(f), C.byref(spl), None, 0) == ERLBM_ERR_ARG
    assert (f.value, spl.value) == (-1, -1)
    assert lib.erlbm_step(None, 1, 0, 0.0) == ERLBM_ERR_ARG
    assert lib.erlbm_set_arith(None, erlbm_amd.ARITH_MOMENT) == ERLBM_ERR_ARG


def test_host_initial_conditions_bitwise(orc):
    import erlbm_amd as E
    for L, u0, tau, poisson, neq in ((64, 2 / 64, 0.62, True, True), (32, 1e-3, 0.8, True, False),
                                      (16, 0.1, 0.55, False, True)):
        assert np.array_equal(E.init_taylor_green(9, L, u0, tau, poisson, neq),
                              orc.init_taylor_green_d2q9(L, u0, tau, poisson, neq))
    assert np.array_equal(E.init_double_shear(128), orc.init_double_shear_d2q9(128))
    assert np.array_equal(E.init_tg_map(16, 0.3), orc.init_tg_map_d3q19(16, 0.3))
    assert np.array_equal(E.obstacle_mask(), orc.obstacle_mask())
    for q in (9, 19):
        assert np.array_equal(E.dist_eq(q, 0.03, -0.02, 0.01 if q == 19 else 0.0, 1.01),
                              orc.dist_eq(q, 0.03, -0.02, 0.01 if q == 19 else 0.0, 1.01))


def test_host_map_axes_bitwise(orc):
    import erlbm_amd as E
    for i, j in ((0, 0), (3, 7), (199, 250), (399, 399)):
        for col, tnh in ((O.COL_EXT_REG, 0.65), (O.COL_REG, 1.0)):
            u, t, tn, n = E.tg_map_point_params(i, j, collision=col, tau_nh=tnh)
            assert (u, t, tn) == orc.tg_map_params(i, j, col=col, tau_nh=tnh)
            assert n == int((100 * 16) / u)
        u, t, tn, n = E.obstacle_map_point_params(i, j)
        assert (u, t, tn) == orc.obstacle_map_params(i, j)
        assert n == int((200 * 12) / u)


def test_errors_are_reported_not_crashing():
    import erlbm_amd as E
    with pytest.raises(E.ErlbmError):
        E.Lattice(7, 4, 4, 1)          # unsupported velocity set: rejected before any HIP call


def test_autotune_candidates_respect_the_31_bit_buffer_limit():
    """The autotuner offers the raw-buffer strip walks (two- to five-step,
    lane-aligned) only where their byte offsets stay below 2^31, as erlbm_step
    does: an FP32-shifted 8192^2 lattice (2.4 GB per buffer) gets none of
    them, a 2048^2 one gets them all.  Host-side; no device call."""
    from erlbm_amd import lattice as Lm
    from erlbm_amd._lib import FP32_SHIFTED, FP64
    walks = {11, 13, 22, 24, 42, 46, 47}
    small = set(Lm.autotune_candidates(9, 2048, 2048, 1, FP32_SHIFTED))
    assert walks <= small, small
    big = set(Lm.autotune_candidates(9, 8192, 8192, 1, FP32_SHIFTED))
    assert not (walks & big), big
    # FP64: the raw-buffer walks stop where 8-byte offsets pass 2^31 as well
    assert {42, 26, 39, 46, 47} <= set(Lm.autotune_candidates(9, 4096, 4096, 1, FP64))
    # small FP64 lattices: the tiles the default rule takes there among them
    assert {1, 3, 7} <= set(Lm.autotune_candidates(9, 256, 256, 1, FP64))
    assert not ({42, 26, 39, 11, 13, 22, 16, 46, 47} & set(Lm.autotune_candidates(9, 8192, 8192, 1, FP64)))
    # D3Q19: the z-walk tiles where the geometry takes them
    assert {33, 31, 30} <= set(Lm.autotune_candidates(19, 256, 256, 256, FP64))
