"""CPU tests of the drop-in boundary: the C-ABI libraries load, export every
symbol the headers in include/ declare, and the host-side logic (options,
grids, argument checks) behaves like the reference's, without a GPU."""
import ctypes
import os
import re

import pytest

from p4pdes_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^#define[^\n]*(\\\n[^\n]*)*", "", text, flags=re.M)   # macros
    names = re.findall(r"^(?!typedef\b)(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*(?:\s+|\s*\*+\s*)([A-Za-z_][A-Za-z_0-9]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


@pytest.mark.parametrize("header", ["p4g.h", "p4g_petsc.h"])
def test_library_exports_every_declared_symbol(header):
    L = capi.lib()
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_fish_library_exports():
    F = capi.fishlib()
    for n in ("fish_run", "fish_session_create", "fish_session_solve", "fish_session_errors",
              "fish_session_solution", "fish_session_solver", "fish_session_destroy",
              "Poisson1DFunctionLocal", "Poisson2DFunctionLocal", "Poisson3DFunctionLocal",
              "Poisson1DJacobianLocal", "Poisson2DJacobianLocal", "Poisson3DJacobianLocal",
              "InitialState"):
        assert hasattr(F, n), n


def test_version():
    assert b"gfx950" in capi.lib().p4g_version()


def test_da_from_options():
    L = capi.lib()
    da = ctypes.c_void_p()
    capi.check(L.p4g_da_create(3, 3, 3, 3, 1.0, 1.0, 1.0, ctypes.byref(da)))
    capi.check(L.p4g_da_set_from_options(da, b"-da_refine 7 -ksp_rtol 1e-10"))
    v = [ctypes.c_int() for _ in range(5)]
    L.p4g_da_get_sizes(da, *[ctypes.byref(x) for x in v])
    assert [x.value for x in v] == [3, 257, 257, 257, 7]
    L.p4g_da_destroy(ctypes.byref(da))
    capi.check(L.p4g_da_create(3, 3, 3, 3, 1.0, 1.0, 1.0, ctypes.byref(da)))
    capi.check(L.p4g_da_set_from_options(da, b"-da_grid_x 5 -da_grid_y 9 -da_refine 1"))
    L.p4g_da_get_sizes(da, *[ctypes.byref(x) for x in v])
    assert [x.value for x in v][:4] == [3, 9, 17, 5]
    L.p4g_da_destroy(ctypes.byref(da))


@pytest.mark.parametrize("opts,fragment", [
    ("-pc_type ilu -pc_factor_levels 1", "ILU(0)"),
    ("-pc_type lu", "lu"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_type ilu -pc_mg_galerkin", "Galerkin"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_sor_omega 1.5", "omega"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_sor_forward", "sor_forward"),
    ("-pc_type mg -pc_mg_levels 3 -p4g_precision mixed", "sor"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_type asm", "asm"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_type jacobi -pc_mg_galerkin -p4g_precision mixed", "galerkin"),
    ("-ksp_type bcgs -pc_type mg", "bcgs"),
    ("-ksp_type gmres -ksp_norm_type unpreconditioned", "left"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_type jacobi -pc_mg_type additive", "additive"),
    ("-pc_type mg -pc_mg_levels 3 -mg_coarse_ksp_type cg", "convergence_test skip"),
    ("-pc_type mg -pc_mg_levels 3 -mg_coarse_ksp_type cg -mg_coarse_ksp_convergence_test skip "
     "-mg_coarse_pc_type sor", "mg_coarse_pc_type sor"),
    ("-pc_type mg -pc_mg_levels 3 -mg_levels_pc_type jacobi -mg_coarse_ksp_type cg "
     "-mg_coarse_ksp_convergence_test skip -mg_coarse_pc_type jacobi -p4g_precision mixed", "FP64"),
    ("-pc_type mg -pc_mg_levels 3 -mg_coarse_ksp_type gmres", "fixed count"),
])
def test_unsupported_options_rejected(opts, fragment):
    L = capi.lib()
    da = ctypes.c_void_p()
    capi.check(L.p4g_da_create(2, 17, 17, 1, 1.0, 1.0, 1.0, ctypes.byref(da)))
    s = ctypes.c_void_p()
    rc = L.p4g_solver_create(da, opts.encode(), ctypes.byref(s))
    assert rc == 56   # PETSC_ERR_SUP
    assert fragment in capi.last_error()
    L.p4g_da_destroy(ctypes.byref(da))


@pytest.mark.parametrize("opts", ["-p4g_fuse_cg maybe", "-p4g_fuse_r 2", "-p4g_fuse_dir on"])
def test_fusion_options_validated(opts):
    L = capi.lib()
    da = ctypes.c_void_p()
    capi.check(L.p4g_da_create(3, 17, 17, 17, 1.0, 1.0, 1.0, ctypes.byref(da)))
    s = ctypes.c_void_p()
    rc = L.p4g_solver_create(da, ("-pc_type mg -mg_levels_pc_type jacobi " + opts).encode(), ctypes.byref(s))
    assert rc == 63   # PETSC_ERR_ARG_OUTOFRANGE
    assert "must be auto, 1 or 0" in capi.last_error()
    L.p4g_da_destroy(ctypes.byref(da))


@pytest.mark.parametrize("args,code", [
    ("-fsh_cx -1 -pc_type mg", 2), ("-fsh_problem manuexp -fsh_cx 2 -pc_type mg", 3),
    ("-fsh_dim 4 -pc_type mg", 1)])
def test_fish_argument_errors(args, code):
    # fish.c:188-193, 214-215: the same codes as the reference, before any GPU work
    rc, text, _, _ = capi.fish_run(args)
    assert rc == code
    assert "PETSC ERROR" in text


def test_no_gpu_fails_loudly():
    if capi.device_count() > 0:
        pytest.skip("a HIP device is present")
    rc, text, _, _ = capi.fish_run("-fsh_dim 2 -da_refine 2 -pc_type mg -mg_levels_pc_type jacobi")
    assert rc != 0 and "no HIP device" in text


def test_grid_shape_matches_oracle():
    from oracle import oracle_py
    for a in ("-fsh_dim 3 -da_refine 3", "-fsh_dim 2 -da_refine 6", "-fsh_dim 1 -da_refine 4 -da_grid_x 5"):
        assert capi.grid_shape(a) == oracle_py.grid_shape(a)
