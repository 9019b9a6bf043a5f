"""The integration guide's double-shear-layer example, compiled as a program.

The ```cpp block under "Example: the double shear layer driver" in
INTEGRATION.md is split into its preprocessor lines and its time loop and
compiled into tests/integration_dsl_example.cpp, which supplies the setup of
src/main_double-shear-layer.cpp:70-154 around it, against include/erlbm.h and
liberlbm.so.

  * CPU: the block compiles and links (every call it makes is declared with
    the argument types it passes);
  * GPU: --L 16 / 32 --t-final 0.125 --dt-out 0.0625 writes k_energy_lbm.dat
    and vel_*.vtk byte-identical to bin/double_shear_layer's and to the
    reference driver's files (tests/golden/ref_outputs.npz).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "erlbm_amd")
LIB = os.path.join(PKG, "liberlbm.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")
HEADING = "## Example: the double shear layer driver"


def example_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index(HEADING)
    m = re.search(r"```cpp\n(.*?)```", text[start:], re.S)
    assert m, "no cpp block under the double shear layer heading"
    lines = m.group(1).splitlines()
    # preprocessor lines (with their continuations) come first, then the loop
    n, cont = 0, False
    while n < len(lines) and (cont or lines[n].startswith("#")):
        cont = lines[n].rstrip().endswith("\\")
        n += 1
    return "\n".join(lines[:n]) + "\n", "\n".join(lines[n:]) + "\n"


def compile_example(out_dir):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (python -m erlbm_amd.build)")
    prelude, body = example_block()
    assert "erlbm_step" in body and "file_energy" in body
    with open(os.path.join(out_dir, "example_prelude.inc"), "w") as fh:
        fh.write(prelude)
    with open(os.path.join(out_dir, "example_body.inc"), "w") as fh:
        fh.write(body)
    exe = os.path.join(out_dir, "integration_dsl_example")
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Werror",
           "-Wno-unused-variable", f"-I{os.path.join(ROOT, 'include')}", f"-I{out_dir}",
           os.path.join(ROOT, "tests", "integration_dsl_example.cpp"), "-o", exe, f"-L{PKG}", "-lerlbm",
           f"-Wl,-rpath,{PKG}"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return exe


def test_example_compiles_and_links(tmp_path):
    compile_example(str(tmp_path))


@pytest.mark.gpu
def test_example_files_equal_driver_and_reference(gpu, golden, tmp_path):
    import numpy as np
    if not os.path.exists(os.path.join(GOLDEN, "ref_outputs.npz")):
        pytest.skip("output fixtures not generated")
    g = golden("ref_outputs.npz")
    names = [str(n) for n in g["dsl_files"]]
    exe = compile_example(str(tmp_path))
    driver = os.path.join(PKG, "bin", "double_shear_layer")
    for L in (16, 32):
        args = ["--L", str(L), "--t-final", "0.125", "--dt-out", "0.0625"]
        ex, dr = tmp_path / f"example{L}", tmp_path / f"driver{L}"
        ex.mkdir()
        dr.mkdir()
        print(f"integration example and driver, L={L}", flush=True)
        for prog, cwd in ((exe, ex), (driver, dr)):
            p = subprocess.run([prog, *args, "--vorticity", "0"], cwd=str(cwd), capture_output=True, timeout=300)
            assert p.returncode == 0, p.stderr.decode()
        want = sorted(n.split("/", 1)[1] for n in names if n.startswith(f"dsl{L}/") and "/vor_" not in n)
        assert sorted(os.listdir(ex)) == want
        assert sorted(os.listdir(dr)) == want
        for fn in want:
            got = (ex / fn).read_bytes()
            assert got == (dr / fn).read_bytes(), (L, fn)
            assert got == bytes(g[f"dsl{L}/{fn}"]), (L, fn)
        assert np.loadtxt(str(ex / "k_energy_lbm.dat")).shape == (3, 2)
