"""The Fortran host (conformal_cubed_sphere_amd/fortran/ccsm.exe): the
reference driver's namelist and step loop (ccsm.F90:2-73) calling the GPU
library through iso_c_binding."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "conformal_cubed_sphere_amd", "fortran", "ccsm.exe")

NAMELIST = """&time_settings
 dt               = {dt}
 run_days         = 0
 run_hours        = {hours}
 run_minutes      = 0
 run_seconds      = 0
 history_interval = {hist}
 integral_scheme  = '{scheme}'
/

&case_select
 case_num = {case}
/

&domain
 mesh_file = 'unused.nc'
 xhalo     = {halo}
 yhalo     = {halo}
/

&dynamic_opt
 reconstruct_scheme = '{recon}'
/

&gpu_opt
{gpu}
/
"""


def _run(tmp_path, timeout=120, **kw):
    args = dict(dt=600.0, hours=6, hist=3600, scheme="RK3_TVD", case=7, halo=3, gpu=" nx = 32", recon="WENO")
    args.update(kw)
    (tmp_path / "namelist.input").write_text(NAMELIST.format(**args))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail("ccsm.exe is not built: run __graft_entry__.build()")


@pytest.mark.parametrize("bad,msg", [
    (dict(scheme="Euler"), "Unknown integral scheme"),
    (dict(gpu=" precision = 'fp64'"), "set nx"),
    (dict(gpu=" nx = 33"), "set nx"),
    (dict(halo=2), "xhalo >= 3"),
    (dict(recon="MUSCL"), "reconstruct_scheme must be"),
])
def test_namelist_validation(tmp_path, bad, msg):
    _need_exe()
    rc, out = _run(tmp_path, **bad)
    assert rc != 0 and msg in out, out


def test_mesh_area_and_cfl_step_before_the_device(tmp_path, ccs):
    """Everything before the first device call runs on the host: the
    generated mesh, the seeded case, and the CFL step (same as the Python
    host's).  Without a GPU the driver must stop loudly at ccs_create."""
    _need_exe()
    rc, out = _run(tmp_path, gpu=" nx = 32\n cfl = 0.5\n seed = 3")
    m = ccs.Mesh(32, halo=3)
    s = ccs.init_case(m, case_num=7, seed=3)
    dt, n = _cfl_with_outputs(ccs, m, s, 6)
    area = re.search(r"Spherical area, numerical, analytical :\s*(\S+)\s+(\S+)", out)
    assert area and abs(float(area.group(1)) / float(np.sum(m.interior(m.areaCell))) - 1) < 1e-13, out
    got = re.search(r"dt from CFL\s+\S+\s*:\s*(\S+)\s+s,\s+(\d+)\s+steps", out)
    assert got and float(got.group(1)) == pytest.approx(dt, rel=1e-15) and int(got.group(2)) == n, out
    if "ccs error in ccs_create" in out:
        assert rc == 2
    else:
        assert rc == 0, out


def _cfl_with_outputs(ccs, m, s, hours):
    # CFL step of the Python host, step count rounded up to whole hours
    _, n = ccs.cfl_time_step(m, s, 0.5, hours * 3600.0)
    n = -(-n // hours) * hours
    return hours * 3600.0 / n, n


def _read_final(path, npad):
    b = path.read_bytes()
    n, h, nsteps, dt = struct.unpack("<iiid", b[:20])
    f = np.frombuffer(b[20:], dtype="<f8")
    assert f.size == 4 * npad
    return n, h, nsteps, dt, {k: f[i * npad:(i + 1) * npad] for i, k in enumerate(("phiG", "u", "v", "phi"))}


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,code", [("RK3_TVD", 0), ("RK4", 1)])
def test_driver_run_matches_oracle_bitwise(tmp_path, ccs, orc, scheme, code):
    """C32, seeded jet with mountain, 6 hours at CFL 0.5 with hourly output:
    the FP64 path through the Fortran host equals the oracle bit for bit, and
    the printed mass change stays at round-off."""
    _need_exe()
    rc, out = _run(tmp_path, scheme=scheme,
                   gpu=" nx = 32\n cfl = 0.5\n seed = 5\n mountain = 1\n final_state_file = 'final.bin'")
    print(out)
    assert rc == 0, out
    m = ccs.Mesh(32, halo=3)
    n, h, nsteps, dt, fin = _read_final(tmp_path / "final.bin", m.npad)
    assert (n, h) == (32, 3)
    s0 = ccs.init_case(m, case_num=7, seed=5, mountain=True)
    dt_py, n_py = _cfl_with_outputs(ccs, m, s0, 6)
    assert dt == dt_py and nsteps == n_py
    o = orc.Oracle(m, s0, dt)
    o.steps(nsteps, scheme=code)
    mask = ~m.corner_mask().ravel()
    for k in ("phiG", "u", "v", "phi"):
        assert np.array_equal(fin[k][mask], o.a[k][mask]), k
    mcr = [float(x.split()[-2]) for x in out.splitlines() if "MCR, ECR" in x]
    assert len(mcr) == 7 and max(abs(x) for x in mcr) < 1e-13
    # history file (io_mod.F90:309-541): initial state + 6 outputs, the last
    # one being the final state
    from scipy.io import netcdf_file

    with netcdf_file(str(tmp_path / "ccsm_output.nc"), "r", mmap=False) as f:
        assert list(f.variables["time"][:]) == list(range(1, 8))
        phi_last = f.variables["phi"][-1].copy().ravel()
        phi_first = f.variables["phi"][0].copy().ravel()
    assert np.array_equal(phi_last, fin["phi"])
    assert np.array_equal(phi_first[mask], s0.phi[mask])


def test_driver_reads_a_mesh_file(tmp_path, ccs):
    """With &domain mesh_file naming a file of the generator's schema, the
    driver takes the grid from it (io_mod.F90:37-135) instead of nx."""
    _need_exe()
    ccs.write_mesh_file(str(tmp_path / "mesh_C16.nc"), 16)
    (tmp_path / "namelist.input").write_text(NAMELIST.format(
        dt=600.0, hours=1, hist=3600, scheme="RK3_TVD", case=2, halo=3, gpu=" precision = 'fp64'",
        recon="WENO").replace(
        "unused.nc", "mesh_C16.nc"))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert "Mesh read from mesh_C16.nc" in out, out
    m = ccs.Mesh(16, halo=3)
    area = re.search(r"Spherical area, numerical, analytical :\s*(\S+)\s+(\S+)", out)
    assert area and abs(float(area.group(1)) / float(np.sum(m.interior(m.areaCell))) - 1) < 1e-12, out
    if "ccs error in ccs_create" in out:
        assert r.returncode == 2
    else:
        assert r.returncode == 0, out


@pytest.mark.gpu
@pytest.mark.parametrize("recon,op", [("4th", "centred4"), ("WENO_ext", "weno_ext")])
def test_driver_operator_variants_match_oracle_bitwise(tmp_path, ccs, orc, recon, op):
    """&dynamic_opt reconstruct_scheme selects the spatial_operators_mod
    variant; 2 hours of C32 through the Fortran host equal the oracle's."""
    _need_exe()
    rc, out = _run(tmp_path, hours=2, recon=recon,
                   gpu=" nx = 32\n cfl = 0.5\n seed = 5\n mountain = 1\n final_state_file = 'final.bin'\n"
                       " history_file = ''")
    print(out)
    assert rc == 0, out
    assert f"reconstruct_scheme {recon}" in out
    m = ccs.Mesh(32, halo=3)
    n, h, nsteps, dt, fin = _read_final(tmp_path / "final.bin", m.npad)
    s0 = ccs.init_case(m, case_num=7, seed=5, mountain=True)
    o = orc.Oracle(m, s0, dt, op=op)
    o.steps(nsteps)
    mask = ~m.corner_mask().ravel()
    for k in ("phiG", "u", "v", "phi"):
        assert np.array_equal(fin[k][mask], o.a[k][mask]), k


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_driver_checkpoint_resume_and_run_record(tmp_path, ccs, precision):
    """final_state_file is a checkpoint: a run of 3 h, then a run of 3 h
    started from it (initial_state_file), ends where one run of 6 h ends, bit
    for bit on the FP64 parity path (every step after the first derives the
    companions from phiG, u, v, as the resumed run's first step does).  The
    FP32 path resumes from the FP64 values of its state, split afresh into
    reference field and deviation, which differs from the uninterrupted run
    at FP32 rounding level; the method amplifies such differences (at C32 to
    ~1e-3 within hours), so there the resumed run is only required to be
    finite and to conserve mass.  The run record is one JSON line with the
    step loop's throughput and the final MCR/ECR."""
    import json

    _need_exe()
    gpu = f" nx = 32\n precision = '{precision}'\n history_file = ''"
    m = ccs.Mesh(32, halo=3)
    npad = m.npad
    full, first, second = tmp_path / "full", tmp_path / "first", tmp_path / "second"
    for d in (full, first, second):
        d.mkdir()
    rc, out = _run(full, hours=6, gpu=gpu + "\n final_state_file = 'end.bin'\n run_record_file = 'run.json'")
    assert rc == 0, out
    rc, out = _run(first, hours=3, gpu=gpu + "\n final_state_file = 'end.bin'")
    assert rc == 0, out
    (second / "start.bin").write_bytes((first / "end.bin").read_bytes())
    rc, out = _run(second, hours=3, gpu=gpu + "\n initial_state_file = 'start.bin'\n final_state_file = 'end.bin'")
    assert rc == 0 and "State read from start.bin" in out, out
    _, _, n_full, _, a = _read_final(full / "end.bin", npad)
    _, _, n_half, _, b = _read_final(second / "end.bin", npad)
    assert n_full == 2 * n_half
    inner = ~m.corner_mask().ravel()
    for k in ("phiG", "u", "v"):
        if precision == "fp64":
            assert np.array_equal(a[k][inner], b[k][inner]), k
        else:
            assert np.all(np.isfinite(b[k][inner])), k
    if precision == "fp32":
        mass = [float(np.sum(m.interior(x["phiG"]))) for x in (a, b)]
        print(f"  total phiG resumed vs uninterrupted: {(mass[1] - mass[0]) / mass[0]:.2e}")
        assert abs(mass[1] - mass[0]) <= 1e-9 * abs(mass[0])
    rec = json.loads((full / "run.json").read_text())
    assert rec["n"] == 32 and rec["halo"] == 3 and rec["precision"] == precision and rec["steps"] == n_full
    assert rec["cell_updates_per_s"] > 0 and abs(rec["mcr"]) < 1e-9 and abs(rec["ecr"]) < 1e-3
    assert json.loads([ln for ln in out.splitlines() if ln.startswith("{")][-1])["steps"] == n_half
