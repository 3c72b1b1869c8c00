"""Write the golden fixtures under tests/golden/ from the reference's own code.

Run in this repository with the reference tree present:

    make -C oracle ref && python tests/golden/make_golden.py

It loads the reference's dynamic-core modules and the mesh generator's map
library as compiled unmodified by `make -C oracle ref` (oracle/_ref/*.so) and
records, for a C6 grid (halo 3, Williamson case 6 on this library's mesh
tables):
  - the inputs: padded metric tables, f, phis, the initial state and its
    phi/uc/vc companions, dt, dx and the extended-lattice G^-1;
  - for each spatial_operators_mod source (default, _2nd, _4th, _6th, _WENO):
    the reference's spatial_operator output and its state after 3 SSP-RK3
    steps (time_scheme_mod.F90:46-115 around the reference operator, as
    oracle/ref_harness.F90 restates it);
  - the generator's xmtoxc (mesh_generation/src/cuco.f90) on the C6 extended
    lattice (13 x 13 points per panel, edges and vertices included).
Output: golden_c6.npz (plain arrays, no pickled objects).
"""
import ctypes
import os
import sys
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
LIBS = {"weno": "libccs_ref.so", "centred2": "libccs_ref_2nd.so", "centred4": "libccs_ref_4th.so",
        "centred6": "libccs_ref_6th.so", "weno_ext": "libccs_ref_WENO.so"}
N, H, CASE, NSTEPS = 6, 3, 6, 3


def p(a):
    return a.ctypes.data_as(_D)


def run(out):
    import conformal_cubed_sphere_amd as ccs

    mesh = ccs.Mesh(N, halo=H)
    s0 = ccs.init_case(mesh, case_num=CASE)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 3600.0)
    ig_ext = mesh.ig_ext()
    for k in ("sqrtG", "G", "IG", "A", "IA", "f", "lon", "lat", "sinlat", "coslat"):
        out["in_" + k] = getattr(mesh, k).copy()
    for k in ("phiG", "u", "v", "phi", "uc", "vc", "phis"):
        out["in_" + k] = getattr(s0, k).copy()
    out["in_ig_ext"] = ig_ext
    out["n"], out["halo"], out["dx"], out["dt"] = N, H, mesh.dx, dt
    for op, name in LIBS.items():
        ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", name))
        ref.ref_init.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_size_t]
        ref.ref_rk3.argtypes = [ctypes.c_int, ctypes.c_double]
        ref.ref_init(N, H, mesh.dx, 64 << 20)
        ref.ref_set_mesh(p(mesh.sqrtG), p(mesh.G), p(mesh.IG), p(mesh.A), p(mesh.IA), p(mesh.f), p(s0.phis),
                         p(mesh.lon), p(mesh.lat), p(mesh.sinlat), p(mesh.coslat))
        ref.ref_set_mesh_ext(p(ig_ext))
        ref.ref_set_state(0, *(p(getattr(s0, k)) for k in ("phiG", "u", "v", "phi", "uc", "vc")))
        t = [np.zeros(mesh.npad) for _ in range(3)]
        ref.ref_spatial_operator(0, *(p(x) for x in t))
        for k, x in zip(("phiG", "u", "v"), t):
            out[f"{op}_tend_{k}"] = mesh.interior(x).copy()
        ref.ref_rk3(NSTEPS, dt)
        st = {k: np.zeros(mesh.npad) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")}
        ref.ref_get_state(0, *(p(st[k]) for k in ("phiG", "u", "v", "phi", "uc", "vc", "zonal", "merid")))
        for k in ("phiG", "u", "v", "phi", "uc", "vc"):
            out[f"{op}_rk3_{k}"] = st[k]
    m = 2 * N + 1
    c = -1.0 + np.arange(m) / N  # mesh_mod.f90:41-47 with dx = 90/N degrees
    P, J, I = np.meshgrid(np.arange(6), np.arange(m), np.arange(m), indexing="ij")
    x1, x2 = np.ascontiguousarray(c[I.ravel()]), np.ascontiguousarray(c[J.ravel()])
    ipan = np.ascontiguousarray(P.ravel() + 1, dtype=np.int32)
    npt = x1.size
    mlib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libccs_ref_mesh.so"))
    xc, d = np.zeros(3 * npt), np.zeros(6 * npt)
    mlib.ref_mesh_points(ctypes.c_int(npt), p(x1), p(x2), ipan.ctypes.data_as(_I), p(xc), p(d))
    out["map_xc"] = xc.reshape(6, m, m, 3)          # [panel, j, i, xyz]
    out["map_dxcdxm"] = d.reshape(6, m, m, 2, 3)    # per map unit, [.., column, row]


def main():
    out, err = {}, []

    def target():
        try:
            run(out)
        except Exception as exc:
            err.append(repr(exc))

    threading.stack_size(512 << 20)  # the reference's automatic arrays live on the stack
    t = threading.Thread(target=target)
    t.start()
    t.join()
    if err:
        raise SystemExit(err[0])
    np.savez_compressed(os.path.join(HERE, "golden_c6.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_c6.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
