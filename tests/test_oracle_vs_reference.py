"""The C oracle against the reference's own Fortran modules.

oracle/_ref/libccs_ref.so is built from the reference sources by
`make -C oracle ref` (when the reference tree is present).  These tests run
on the CPU: the oracle must reproduce the reference's WENO_limiter, initial
conditions, spatial_operator, halo fill and RK3 steps bit for bit, and its
diagnostics exactly.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libccs_ref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB),
                                reason="reference modules not built (oracle/_ref needs the reference tree)")


def _compare(n, h, case_num, nsteps, plant=False, op="weno"):
    r = subprocess.run([sys.executable, os.path.join(HERE, "ref_compare.py"), str(n), str(h), str(case_num),
                        str(nsteps), f"op={op}"] + (["plant"] if plant else []), capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    line = next(x for x in r.stdout.splitlines() if x.startswith("{"))  # the reference also prints
    out = json.loads(line)
    assert "error" not in out, out["error"]
    return out


@pytest.mark.parametrize("n,h,case_num,nsteps", [(16, 3, 2, 20), (16, 4, 5, 10), (24, 3, 6, 10), (32, 3, 6, 5)])
def test_oracle_reproduces_reference_modules(n, h, case_num, nsteps):
    out = _compare(n, h, case_num, nsteps)
    print(json.dumps(out))
    assert out["weno_mismatches"] == 0
    assert all(v <= 4 * 2.0 ** -52 for v in out["init_max_rel_diff"].values()), out["init_max_rel_diff"]
    assert out["init_phis_max_abs_diff"] == 0.0
    assert all(out["tend_bitwise_equal"]), out["tend_max_rel_diff"]
    assert out["rk3_bitwise_equal"], out["rk3_max_rel_diff"]
    assert out["diag_mass"][0] == out["diag_mass"][1]
    assert out["diag_energy"][0] == out["diag_energy"][1]


def test_comparison_detects_one_ulp_planted_in_the_oracle():
    # the RK3 comparison is not vacuous: one ulp in one cell of the oracle's
    # initial phiG must show up after two steps
    out = _compare(16, 3, 2, 2, plant=True)
    assert not out["rk3_bitwise_equal"]
    assert 0.0 < out["rk3_max_rel_diff"]["phiG"] < 1e-12


VARIANT_LIBS = {"centred2": "libccs_ref_2nd.so", "centred4": "libccs_ref_4th.so", "centred6": "libccs_ref_6th.so",
                "weno_ext": "libccs_ref_WENO.so"}


@pytest.mark.parametrize("op", sorted(VARIANT_LIBS))
@pytest.mark.parametrize("n,h,case_num,nsteps", [(16, 3, 2, 10), (24, 4, 5, 5)])
def test_oracle_operator_variants_reproduce_reference_modules(op, n, h, case_num, nsteps):
    # the reference built with spatial_operators_mod_{2nd,4th,6th,WENO}.F90 in
    # place of spatial_operators_mod.F90, against the oracle's operator variant
    lib = os.path.join(ROOT, "oracle", "_ref", VARIANT_LIBS[op])
    if not os.path.exists(lib):
        pytest.skip("reference variant module not built")
    out = _compare(n, h, case_num, nsteps, op=op)
    print(json.dumps(out))
    assert out["op"] == op
    assert all(out["tend_bitwise_equal"]), out["tend_max_rel_diff"]
    assert out["rk3_bitwise_equal"], out["rk3_max_rel_diff"]
    assert out["diag_mass"][0] == out["diag_mass"][1]
    assert out["diag_energy"][0] == out["diag_energy"][1]


def test_variant_comparison_detects_one_ulp_planted_in_the_oracle():
    out = _compare(16, 3, 2, 2, plant=True, op="weno_ext")
    assert not out["rk3_bitwise_equal"]
