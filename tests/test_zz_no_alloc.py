"""ccs_diagnostics_noalloc: mass and energy of a state in caller-owned device
memory, with the library allocating nothing (its partial sums sit in a
file-scope device array).  Called through the C ABI on device arrays that the
caller owns (torch tensors as plain device memory), twice in a row, the second
time on a smaller grid, and compared bit for bit with ccs_diagnostics on the
same state, which runs the same kernels on the handle's own buffers.  The
test plays the caller: it owns its device memory through the HIP runtime the
library itself uses (hipMalloc / hipMemcpy / hipFree by ctypes)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(ccs, n, seed):
    mesh = ccs.Mesh(n, halo=3, radius=6.371e6)
    s0 = ccs.init_case(mesh, case_num=7, seed=seed, mountain=True)
    dt, _ = ccs.cfl_time_step(mesh, s0, 0.5, 6 * 3600.0)
    with ccs.DynamicalCore(mesh, dt, precision="fp64") as core:
        core.set_state(s0)
        core.step(5)
        st = core.get_state()
        ref = core.diagnostics()
    return mesh, st, ref


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime libccs.so is linked against (already loaded)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def test_noalloc_diagnostics_equal_handle_diagnostics(ccs):
    from conformal_cubed_sphere_amd._lib import lib

    hip = _hip()
    h2d, d2h = 1, 2
    for n, seed in ((48, 3), (24, 4)):  # the second call on a smaller grid
        mesh, st, (mass, energy) = _state(ccs, n, seed)
        ig = mesh.IG.reshape(-1, 4)  # per cell (1,1), (2,1), (1,2), (2,2)
        arrays = [np.ascontiguousarray(a, dtype=np.float64)
                  for a in (st.phiG, st.u, st.v, mesh.sqrtG, ig[:, 0], ig[:, 2], ig[:, 1], ig[:, 3], st.phis)]
        ptrs = []
        try:
            for a in arrays + [np.full(2, np.nan)]:
                p = ctypes.c_void_p()
                assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
                ptrs.append(p)
                assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, h2d) == 0
            err = ctypes.create_string_buffer(256)
            print(f"C{n}: ccs_diagnostics_noalloc on caller device arrays", flush=True)
            rc = lib.ccs_diagnostics_noalloc(n, mesh.halo, *ptrs, None, err, 256)
            assert rc == 0, err.value.decode()
            assert hip.hipDeviceSynchronize() == 0
            got = np.zeros(2)
            assert hip.hipMemcpy(got.ctypes.data, ptrs[-1], got.nbytes, d2h) == 0
        finally:
            for p in ptrs:
                hip.hipFree(p)
        print(f"  mass {got[0]:.17e} / {mass:.17e}, energy {got[1]:.17e} / {energy:.17e}", flush=True)
        assert got[0] == mass and got[1] == energy


def test_noalloc_rejects_bad_sizes(ccs):
    from conformal_cubed_sphere_amd._lib import lib

    err = ctypes.create_string_buffer(256)
    rc = lib.ccs_diagnostics_noalloc(20000, 3, *([None] * 10), None, err, 256)
    assert rc == 1 and b"maximum" in err.value
    rc = lib.ccs_diagnostics_noalloc(8, 2, *([None] * 10), None, err, 256)
    assert rc == 1
