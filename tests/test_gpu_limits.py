"""Size limits of the multi-step walks.  Their raw buffers take 31-bit byte
offsets, so a lattice whose FP64 population buffer stays below 2^31 bytes
runs the five-step walk, and a larger one the two-step 64x8 tiles (64-bit
addressing); both bitwise the strict oracle (D2Q9 double shear, E-RLBM).

  * L = 5120: 26,214,400 nodes, 1.89 GB per buffer — the default
    five-step walk, 11 steps (the one-step first step, two five-step
    launches);
  * L = 5632: 31,719,424 nodes, 2.28 GB per buffer — no strip walk fits,
    the 64x8 tiles, 5 steps (the one-step first step, two two-step
    launches).
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

NTHREADS = min(16, len(os.sched_getaffinity(0)))


@pytest.mark.parametrize("L,n,want_spl", [(5120, 11, 5), (5632, 5, 2)])
def test_largest_lattices_bitwise(gpu, orc, L, n, want_spl):
    tau = 0.5 + 3.0 * (1.0 / 32.0) * L / 80000.0
    f0 = gpu.init_double_shear(L)
    ini, st, nf = orc.stream_map(9, L, L, 1)
    print(f"oracle: DSL L={L}, {n} steps", flush=True)
    want = orc.run_steps(9, f0, st, n, O.COL_EXT_REG, tau, 0.65, nthreads=NTHREADS)
    with gpu.Lattice(9, L, L, 1) as lat:
        lat.set_params(tau, 0.65, 1.0, O.COL_EXT_REG)
        lat.upload(f0)
        shape, spl, desc = lat.step_shape()
        print(f"L={L}: {desc} ({spl} steps per launch)", flush=True)
        assert spl == want_spl
        lat.step(n)
        assert lat.status() == (n, 0)
        got = lat.download()
    assert np.array_equal(got, want)
