"""GPU test of p4g_pc_apply_device_noalloc, the variant of PCApply
(p4g_pc_apply_device) that allocates nothing: its cycle's work arrays come
from one device array the library holds per GPU.  Through the C ABI it gives
the existing entry point's result bit for bit on two calls in a row, the
second on a smaller hierarchy (so the scratch holds the first call's larger
contents), and refuses an un-set-up solver and a hierarchy beyond its size."""
import ctypes

import numpy as np
import pytest

from p4pdes_amd import capi

pytestmark = pytest.mark.gpu

MG = "-pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"
# the first hierarchy larger than the second; V(2,2) and V(1,1), 3D and 2D
CASES = [("-fsh_dim 3 -fsh_problem manusin -da_refine 5 %s -mg_levels_ksp_max_it 2" % MG, 1),
         ("-fsh_dim 2 -fsh_problem manusin -da_refine 6 %s -mg_levels_ksp_max_it 1" % MG, 2)]


def _to_device(x):
    L = capi.lib()
    p = L.p4g_device_malloc(x.nbytes)
    assert p
    capi.check(L.p4g_memcpy_htod(p, x.ctypes.data, x.nbytes))
    return p


def _run(fn, solver, r):
    L = capi.lib()
    dr = _to_device(r)
    dz = L.p4g_device_malloc(r.nbytes)
    assert dz
    try:
        capi.check(fn(solver, dr, dz))
        z = np.empty_like(r)
        capi.check(L.p4g_memcpy_dtoh(z.ctypes.data, dz, r.nbytes))
    finally:
        L.p4g_device_free(dr)
        L.p4g_device_free(dz)
    return z


def test_noalloc_matches_pc_apply_two_calls():
    L = capi.lib()
    sessions = []
    try:
        refs = []
        for args, seed in CASES:
            print("noalloc case", args, flush=True)
            s = capi.Session(args)
            sessions.append(s)
            capi.check(L.p4g_solver_setup(s.solver))
            r = np.random.default_rng(seed).standard_normal(s.n)
            refs.append((r, _run(L.p4g_pc_apply_device, s.solver, r)))
        # two no-allocation calls in a row: the larger hierarchy, then the smaller
        for (args, _), s, (r, zref) in zip(CASES, sessions, refs):
            z = _run(L.p4g_pc_apply_device_noalloc, s.solver, r)
            assert np.all(np.isfinite(z))
            assert np.array_equal(z, zref), (args, float(np.max(np.abs(z - zref))))
        # and the existing entry point is unchanged after them
        s, (r, zref) = sessions[0], refs[0]
        assert np.array_equal(_run(L.p4g_pc_apply_device, s.solver, r), zref)
    finally:
        for s in sessions:
            s.close()


def test_noalloc_refuses_unset_solver_and_oversize_hierarchy():
    L = capi.lib()
    da = ctypes.c_void_p()
    capi.check(L.p4g_da_create(3, 9, 9, 9, 1.0, 1.0, 1.0, ctypes.byref(da)))
    sv = ctypes.c_void_p()
    capi.check(L.p4g_solver_create(da, b"-pc_type mg -mg_levels_pc_type jacobi", ctypes.byref(sv)))
    x = L.p4g_device_malloc(8 * 729)
    try:
        rc = L.p4g_pc_apply_device_noalloc(sv, x, x)
        assert rc == 73 and "set the solver up first" in capi.last_error()   # PETSC_ERR_ORDER
    finally:
        L.p4g_device_free(x)
        L.p4g_solver_destroy(ctypes.byref(sv))
        L.p4g_da_destroy(ctypes.byref(da))
    args = "-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -mg_levels_ksp_max_it 1" % MG   # 257^3
    print("noalloc oversize", args, flush=True)
    with capi.Session(args) as s:
        capi.check(L.p4g_solver_setup(s.solver))
        d = L.p4g_device_malloc(8 * s.n)
        try:
            rc = L.p4g_pc_apply_device_noalloc(s.solver, d, d)
            assert rc == 63 and "more than the 128 MiB scratch" in capi.last_error()   # PETSC_ERR_ARG_OUTOFRANGE
        finally:
            L.p4g_device_free(d)
