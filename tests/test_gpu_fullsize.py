"""The double-shear-layer workload at its full size: D2Q9 E-RLBM, periodic
L = 2048 (4,194,304 nodes), tau = 0.5 + 3 U0 L / Re (Re = 80000, U0 = 1/32),
tau_nh = 0.65, equilibrium initialisation of src/main_double-shear-layer.cpp:
48-66, 1000 steps, as the default kernels run it (five steps per launch in
FP64, four in FP32-shifted storage), and on the three- and four-step
walks.

  * FP64: bitwise the strict oracle (the reference's operation order, no FMA
    contraction); max relative population error <= 1e-10 against the
    FMA-contracted oracle (the reference's default build);
  * FP64 with the moment-space collision arithmetic (ERLBM_ARITH_MOMENT):
    max relative population error <= 1e-10 against both oracle builds;
  * FP32-shifted storage: bitwise the one-step FP32 path, and relative L2
    velocity error against the FP64 oracle below 5e-6 (the storage format's
    error budget at this size and step count).

The oracle runs once per module (about 25 s on 16 host threads per build).
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

L, N = 2048, 1000
TAU = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
TAU_NH = 0.65
FP32_L2_BOUND = 5e-6
NTHREADS = min(16, len(os.sched_getaffinity(0)))


@pytest.fixture(scope="module")
def dsl_setup(orc):
    import erlbm_amd as E
    f0 = E.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    return f0, st


@pytest.fixture(scope="module")
def strict_1000(orc, dsl_setup):
    f0, st = dsl_setup
    print(f"oracle (strict): DSL L={L}, {N} steps, {NTHREADS} threads", flush=True)
    return orc.run_steps(9, f0, st, N, O.COL_EXT_REG, TAU, TAU_NH, nthreads=NTHREADS)


@pytest.fixture(scope="module")
def fma_1000(dsl_setup):
    f0, st = dsl_setup
    print(f"oracle (FMA): DSL L={L}, {N} steps, {NTHREADS} threads", flush=True)
    return O.Oracle(fma=True).run_steps(9, f0, st, N, O.COL_EXT_REG, TAU, TAU_NH, nthreads=NTHREADS)


def _gpu_run(gpu, f0, precision, arith=0, calls=(N,)):
    with gpu.Lattice(9, L, L, 1, precision=precision) as lat:
        lat.set_params(TAU, TAU_NH, 1.0, O.COL_EXT_REG)
        lat.set_arith(arith)
        lat.upload(f0)
        shape, spl, desc = lat.step_shape()
        for n in calls:
            lat.step(n)
        assert lat.status() == (N, 0)
        return lat.download(), (shape, spl, desc)


def _vel(f):
    c = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [1, -1], [-1, 1]], float)
    return (f @ c) / f.sum(1)[:, None]


@pytest.mark.parametrize("fuse,want_spl", [(None, 5), ("46", 4), ("42", 3)])
def test_dsl2048_1000_steps_default_kernel_bitwise(gpu, monkeypatch, dsl_setup, strict_1000, fuse, want_spl):
    """The default kernel (five steps per launch) and the four- and
    three-step lane-aligned walks."""
    f0, _ = dsl_setup
    if fuse:
        monkeypatch.setenv("ERLBM_FUSE", fuse)
    got, (shape, spl, desc) = _gpu_run(gpu, f0, gpu.FP64)
    print(f"GPU kernel: {desc} (shape {shape}, {spl} steps per launch)", flush=True)
    assert spl == want_spl
    assert np.array_equal(got, strict_1000)


def test_dsl2048_1000_steps_within_1e10_of_fma_oracle(gpu, dsl_setup, fma_1000):
    f0, st = dsl_setup
    got, _ = _gpu_run(gpu, f0, gpu.FP64)
    err = float(np.max(np.abs(got - fma_1000) / np.abs(fma_1000)))
    print(f"max relative population error vs the FMA build: {err:.3e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("fuse,want_spl", [(None, 5), ("46", 4), ("42", 3)])
def test_dsl2048_moment_arith_1000_steps_within_1e10(gpu, monkeypatch, dsl_setup, strict_1000, fma_1000, fuse,
                                                     want_spl):
    f0, _ = dsl_setup
    if fuse:
        monkeypatch.setenv("ERLBM_FUSE", fuse)
    got, (shape, spl, desc) = _gpu_run(gpu, f0, gpu.FP64, gpu.ARITH_MOMENT)
    print(f"GPU kernel, moment-space arithmetic: {desc} (shape {shape}, {spl} steps per launch)", flush=True)
    assert spl == want_spl
    e_strict = float(np.max(np.abs(got - strict_1000) / np.abs(strict_1000)))
    e_fma = float(np.max(np.abs(got - fma_1000) / np.abs(fma_1000)))
    print(f"max relative population error: {e_strict:.3e} vs strict, {e_fma:.3e} vs FMA build", flush=True)
    assert 0.0 < e_strict <= 1e-10
    assert e_fma <= 1e-10


def test_dsl2048_moment_arith_call_pattern(gpu, dsl_setup, strict_1000):
    """The same 1000 steps in calls of 500, 1, 253 and 246 steps: one-,
    three- and four-step remainders (the one-step kernel in the reference's
    order, the three- and four-step lane-aligned walks) between five-step
    launches."""
    f0, _ = dsl_setup
    got, _ = _gpu_run(gpu, f0, gpu.FP64, gpu.ARITH_MOMENT, calls=(500, 1, 253, 246))
    err = float(np.max(np.abs(got - strict_1000) / np.abs(strict_1000)))
    print(f"calls 500+1+253+246: max relative population error {err:.3e} vs strict", flush=True)
    assert 0.0 < err <= 1e-10


def test_dsl2048_fp32_shifted_moment_arith_1000_steps(gpu, dsl_setup, strict_1000):
    """FP32-shifted storage with the moment-space collision (FP64
    arithmetic): the storage format's error budget against the FP64 oracle
    holds, and the run differs from the exact-arithmetic FP32 run only at
    FP32 rounding level."""
    f0, _ = dsl_setup
    got, (shape, spl, desc) = _gpu_run(gpu, f0, gpu.FP32_SHIFTED, gpu.ARITH_MOMENT)
    print(f"GPU FP32 kernel, moment-space arithmetic: {desc} (shape {shape}, {spl} steps per launch)", flush=True)
    assert spl == 5
    exact, _ = _gpu_run(gpu, f0, gpu.FP32_SHIFTED)
    u, uw, ue = _vel(got), _vel(strict_1000), _vel(exact)
    err = float(np.sqrt(np.sum((u - uw) ** 2) / np.sum(uw ** 2)))
    diff = float(np.sqrt(np.sum((u - ue) ** 2) / np.sum(ue ** 2)))
    mass = abs(got.sum() - strict_1000.sum()) / strict_1000.sum()
    print(f"FP32 moment vs FP64 oracle: L2 velocity error {err:.3e}; vs FP32 exact arithmetic {diff:.3e}; "
          f"mass drift {mass:.2e}", flush=True)
    assert not np.array_equal(got, exact)
    assert err < FP32_L2_BOUND
    assert diff < FP32_L2_BOUND
    assert mass < 1e-7


def test_dsl2048_fp32_shifted_1000_steps(gpu, monkeypatch, dsl_setup, strict_1000):
    f0, _ = dsl_setup
    got, (shape, spl, desc) = _gpu_run(gpu, f0, gpu.FP32_SHIFTED)
    print(f"GPU FP32 default kernel: {desc} (shape {shape}, {spl} steps per launch)", flush=True)
    assert spl == 4
    monkeypatch.setenv("ERLBM_FUSE", "0")
    one, (shape1, spl1, _) = _gpu_run(gpu, f0, gpu.FP32_SHIFTED)
    assert (shape1, spl1) == (0, 1)
    assert np.array_equal(got, one)
    u, uw = _vel(got), _vel(strict_1000)
    err = float(np.sqrt(np.sum((u - uw) ** 2) / np.sum(uw ** 2)))
    mass = abs(got.sum() - strict_1000.sum()) / strict_1000.sum()
    print(f"FP32-shifted vs FP64 oracle: L2 velocity error {err:.3e}, mass drift {mass:.2e}", flush=True)
    assert err < FP32_L2_BOUND
    assert mass < 1e-7
