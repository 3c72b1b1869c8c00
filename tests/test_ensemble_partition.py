"""The persistent ensemble kernel splits the global tile list into group
ranges with ccs::ens_group (csrc/ccs_share.h), whose workgroups claim tiles
from a counter per group.  Compile that header on the CPU and check, for
every grid size 1..80 (grids smaller than the 8 XCD groups included) and
tile counts 1..300, that the ranges cover every tile once and that a
simulated run of claims processes every tile exactly once."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "conformal_cubed_sphere_amd", "csrc")


def test_partition_covers_every_tile_once(tmp_path):
    exe = tmp_path / "share_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, os.path.join(HERE, "share_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("OK")
