"""Compare the C oracle with the reference's own dynamic-core modules.

Test infrastructure.  Loads oracle/_ref/libccs_ref.so (the reference's
constants/parameters/mesh/stat/tend/projection/ghost/spatial_operators/diag/
test_case modules compiled from their sources, see oracle/Makefile `ref`)
and oracle/build/libccs_oracle.so, runs both on the same product mesh, and
prints one JSON line of comparison results.  With op=NAME the reference library
built with that spatial operator source (libccs_ref_{2nd,4th,6th,WENO}.so) is
compared with the oracle's matching operator.  Run in its own process
(`python tests/ref_compare.py N HALO CASE NSTEPS [plant] [op=NAME]`): the reference allocates
its automatic arrays on the stack, so the work runs on a thread with a large
stack, and each process initialises the reference modules once.
"""
import ctypes
import json
import os
import sys
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libccs_ref.so")
# oracle operator name -> reference library built with that operator source
REF_LIBS = {"weno": "libccs_ref.so", "centred2": "libccs_ref_2nd.so", "centred4": "libccs_ref_4th.so",
            "centred6": "libccs_ref_6th.so", "weno_ext": "libccs_ref_WENO.so"}

_D = ctypes.POINTER(ctypes.c_double)


def p(a):
    return a.ctypes.data_as(_D)


def run(n, h, case_num, nsteps, out, plant=False, op="weno"):
    import conformal_cubed_sphere_amd as ccs
    import oracle_lib

    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", REF_LIBS[op]))
    ref.ref_init.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_size_t]
    has_weno = op in ("weno", "weno_ext")
    if has_weno:
        ref.ref_weno.restype = ctypes.c_double
        ref.ref_weno.argtypes = [_D, ctypes.c_int]
    ref.ref_rk3.argtypes = [ctypes.c_int, ctypes.c_double]
    ref.ref_diag.argtypes = [ctypes.c_int, _D, _D]

    mesh = ccs.Mesh(n, halo=h)  # radius 6371229 = constants_mod.F90:11
    ref.ref_init(n, h, mesh.dx, 64 << 20)
    zero = np.zeros(mesh.npad)
    ref.ref_set_mesh(p(mesh.sqrtG), p(mesh.G), p(mesh.IG), p(mesh.A), p(mesh.IA), p(mesh.f), p(zero),
                     p(mesh.lon), p(mesh.lat), p(mesh.sinlat), p(mesh.coslat))
    inner = ~mesh.corner_mask().ravel()
    res = {"op": op}
    ig_ext = mesh.ig_ext() if op == "weno_ext" else None
    if ig_ext is not None:
        ref.ref_set_mesh_ext(p(ig_ext))

    # WENO_limiter on random windows of several scales, both directions
    rng = np.random.default_rng(n * 100 + case_num)
    worst = 0
    for _ in range(4000 if has_weno else 0):
        q = rng.normal(size=5) * 10.0 ** rng.integers(-4, 19)
        for d in (1, -1):
            a = ref.ref_weno(p(q), d)
            b = oracle_lib.weno(q, d)
            worst = max(worst, 0 if a == b else 1)
    res["weno_mismatches"] = worst

    # initial state: reference initTestCase vs product vs oracle
    ref.ref_init_case(case_num)
    st = {k: np.zeros(mesh.npad) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")}
    ref.ref_get_state(0, *(p(st[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")))
    phis = np.zeros(mesh.npad)
    ref.ref_get_phis(p(phis))
    prod = ccs.init_case(mesh, case_num=case_num)
    # initial states built with libm transcendentals: the reference's loops
    # may be vectorised with a different sin/cos, so compare in ulps
    res["init_max_rel_diff"] = {k: float(np.max(np.abs(getattr(prod, k)[inner] - st[k][inner]) /
                                                np.maximum(np.abs(st[k][inner]), 1e-300)))
                                for k in ("phiG", "u", "v", "phi", "uc", "vc")}
    res["init_phis_max_abs_diff"] = float(np.max(np.abs(prod.phis[inner] - phis[inner])))

    # from here on both sides start from the identical (product) state
    ref.ref_set_mesh(p(mesh.sqrtG), p(mesh.G), p(mesh.IG), p(mesh.A), p(mesh.IA), p(mesh.f), p(prod.phis),
                     p(mesh.lon), p(mesh.lat), p(mesh.sinlat), p(mesh.coslat))
    ref.ref_set_state(0, *(p(getattr(prod, k)) for k in ("phiG", "u", "v", "phi", "uc", "vc")))

    # tendency of the initial state
    orc = oracle_lib.Oracle(mesh, prod, 60.0, op=op, ig_ext=ig_ext)
    to = orc.tendency()
    tr = [np.zeros(mesh.npad) for _ in range(3)]
    ref.ref_spatial_operator(0, *(p(x) for x in tr))
    res["tend_max_rel_diff"] = [float(np.max(np.abs(mesh.interior(a) - b)) / np.max(np.abs(b)))
                                for a, b in zip(tr, to)]
    res["tend_bitwise_equal"] = [bool(np.array_equal(mesh.interior(a), b)) for a, b in zip(tr, to)]

    # nsteps of RK3 (reference spatial operator + ghost pieces) vs the oracle
    dt, _ = ccs.cfl_time_step(mesh, prod, 0.5, 86400.0)
    orc = oracle_lib.Oracle(mesh, prod, dt, op=op, ig_ext=ig_ext)
    if plant:  # one ulp in one interior cell of the oracle's input only
        d = (n // 2 + mesh.halo) + mesh.P * ((n // 3 + mesh.halo) + mesh.P * 2)
        orc.a["phiG"][d] = np.nextafter(orc.a["phiG"][d], np.inf)
    orc.steps(nsteps)
    ref.ref_rk3(nsteps, dt)
    ref.ref_get_state(0, *(p(st[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")))
    res["rk3_max_rel_diff"] = {k: float(np.max(np.abs(st[k][inner] - orc.a[k][inner])) /
                                        np.max(np.abs(orc.a[k][inner]))) for k in ("phiG", "u", "v", "phi", "uc", "vc")}
    res["rk3_bitwise_equal"] = all(bool(np.array_equal(st[k][inner], orc.a[k][inner]))
                                   for k in ("phiG", "u", "v", "phi", "uc", "vc"))
    m, e = ctypes.c_double(), ctypes.c_double()
    ref.ref_diag(0, ctypes.byref(m), ctypes.byref(e))
    om, oe, _, _ = orc.diagnostics()
    res["diag_mass"] = [m.value, om]
    res["diag_energy"] = [e.value, oe]
    res["dt"] = dt
    out.update(res)


def main():
    n, h, case_num, nsteps = (int(x) for x in sys.argv[1:5])
    plant = "plant" in sys.argv[5:]
    op = next((x[3:] for x in sys.argv[5:] if x.startswith("op=")), "weno")
    out = {}
    err = []

    def target():
        try:
            run(n, h, case_num, nsteps, out, plant, op)
        except Exception as exc:  # report in the JSON line
            err.append(repr(exc))

    threading.stack_size(512 << 20)
    t = threading.Thread(target=target)
    t.start()
    t.join()
    if err:
        out["error"] = err[0]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
