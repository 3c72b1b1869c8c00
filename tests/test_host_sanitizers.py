"""The library's host code, and the CPU oracle that checks it, under
AddressSanitizer and UndefinedBehaviorSanitizer.

ccs_mesh.cpp, ccs_cases.cpp and ccs_netcdf.cpp hold no device code; they are
compiled here with g++ -fsanitize=address,undefined together with
tests/host_sanitize_driver.cpp, which runs the mesh map, metric tables,
initial cases, extended-lattice metric, halo scatter rules and the netCDF
mesh/history files at several sizes, and feeds truncated mesh files to the
reader.  Any sanitizer report makes the driver exit nonzero.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "conformal_cubed_sphere_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_sanitize_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(HERE, "host_sanitize_driver.cpp"), os.path.join(CSRC, "ccs_mesh.cpp"),
           os.path.join(CSRC, "ccs_cases.cpp"), os.path.join(CSRC, "ccs_netcdf.cpp"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    # verify_asan_link_order=0: the sanitizer runtime need not be the first
    # library in the process (a preloaded library may come before it)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "0 failures" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc and g++")
def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    # the checker itself: metrics, cases 2/5/6/7, all five operator sources,
    # RK3 and RK4 steps, halo fill and diagnostics at n = 4, 8, 16, halo 3/4
    san = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-ffp-contract=off", "-fopenmp"]
    obj = str(tmp_path / "ccs_oracle.o")
    c = subprocess.run(["gcc", "-std=gnu11", *san, "-c", os.path.join(ROOT, "oracle", "ccs_oracle.c"), "-o", obj],
                       capture_output=True, text=True, timeout=600)
    assert c.returncode == 0, c.stderr[-3000:]
    exe = str(tmp_path / "oracle_sanitize_driver")
    b = subprocess.run(["g++", "-std=c++17", *san, "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "oracle"), "-I" + CSRC,
                        os.path.join(HERE, "oracle_sanitize_driver.cpp"), os.path.join(CSRC, "ccs_mesh.cpp"), obj,
                        "-o", exe], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "0 failures" in r.stdout
