"""conformal_cubed_sphere_amd/examples/c_host.c: a plain C host of the time
step through include/ccs.h alone.  The CPU test compiles it with -Wall
-Wextra -Werror and links it against libccs.so; the GPU test runs it in both
precisions and checks its mass budget."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "conformal_cubed_sphere_amd")


def _build(tmp_path):
    exe = str(tmp_path / "c_host")
    r = subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(PKG, "examples", "c_host.c"), "-L" + PKG, "-lccs", "-Wl,-rpath," + PKG, "-lm",
                        "-o", exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_c_host_example_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libccs.so")):
        pytest.skip("libccs.so not built")
    _build(tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bound", [("fp32", 1e-9), ("fp64", 1e-12)])
def test_c_host_example_runs(tmp_path, precision, bound):
    exe = _build(tmp_path)
    print(f"c_host 48 {precision} 50", flush=True)
    r = subprocess.run([exe, "48", precision, "50"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("ok"), r.stderr
    dm = float(re.search(r"mass_change=(\S+)", r.stdout).group(1))
    assert abs(dm) <= bound
