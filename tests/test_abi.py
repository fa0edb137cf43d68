"""CPU tests of the drop-in boundary: the C-ABI libraries load, export every
symbol the headers in include/ declare, and the host-side logic (options,
grids, argument checks) behaves like the reference's, without a GPU."""
import ctypes
import os
import re

import numpy as np
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


# PETSc's KSPConvergedReason values (petscksp.h, PETSc >= 3.21) for the
# reasons the solvers report
PETSC_KSP_REASONS = {"CONVERGED_RTOL": 2, "CONVERGED_ATOL": 3, "CONVERGED_ITS": 4, "DIVERGED_NULL": -2,
                     "DIVERGED_ITS": -3, "DIVERGED_DTOL": -4, "DIVERGED_BREAKDOWN": -5,
                     "DIVERGED_INDEFINITE_PC": -8, "DIVERGED_NANORINF": -9, "DIVERGED_INDEFINITE_MAT": -10}


def test_reason_codes_follow_petsc():
    text = open(os.path.join(ROOT, "include", "p4g.h")).read()
    macros = {m: int(v) for m, v in re.findall(r"^#define P4G_((?:CONVERGED|DIVERGED)_[A-Z_]+) \(?(-?\d+)\)?",
                                                 text, flags=re.M)}
    assert macros == PETSC_KSP_REASONS
    oracle = open(os.path.join(ROOT, "oracle", "oracle_internal.h")).read()
    omacros = {m: int(v) for m, v in re.findall(r"^#define R_((?:CONVERGED|DIVERGED)_[A-Z_]+) \(?(-?\d+)\)?",
                                                  oracle, flags=re.M)}
    assert omacros == PETSC_KSP_REASONS
    # the Python names of the reported reasons
    assert all(PETSC_KSP_REASONS[name] == code for code, name in capi.REASONS.items())


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


@pytest.mark.parametrize("opts,fragment", [
    ("-p4g_x_defer 16", "-p4g_x_defer must be in 0..15"), ("-p4g_x_defer -2", "-p4g_x_defer must be in 0..15"),
    ("-p4g_rr_rows 6", "-p4g_rr_rows must be 4 or 8"), ("-p4g_rr_planes 65", "-p4g_rr_planes must be in 0..64"),
    ("-p4g_post_fuse 3", "must be auto, 1 or 0"), ("-p4g_lds yes2", "must be auto, 1 or 0")])
def test_kernel_form_options_validated(opts, fragment):
    # the device options that choose kernel forms: out-of-range values are
    # rejected at creation with PETSC_ERR_ARG_OUTOFRANGE and a message
    L = capi.lib()
    da = ctypes.c_void_p()
    capi.check(L.p4g_da_create(3, 17, 17, 17, 1.0, 1.0, 1.0, ctypes.byref(da)))
    s = ctypes.c_void_p()
    rc = L.p4g_solver_create(da, ("-pc_type mg -mg_levels_pc_type jacobi " + opts).encode(), ctypes.byref(s))
    assert rc == 63, capi.last_error()   # PETSC_ERR_ARG_OUTOFRANGE
    assert fragment in capi.last_error()
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


REFERENCE_CH6 = "/root/reference/c/ch6"


def _cc(args, cwd):
    import subprocess
    r = subprocess.run(["gcc"] + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE_CH6, "poissonfunctions.c")),
                    reason="the reference tree is not present on this machine")
def test_reference_callbacks_compile_unchanged_and_match(tmp_path):
    """The drop-in claim: the reference's own c/ch6/poissonfunctions.c, compiled
    unchanged against include/petsc.h (-std=c99 -Wall -Werror), registers with
    p4g_da_set_function_local / p4g_da_set_jacobian_local, and defines the same
    residual and the same operator (written by p4g_da_mat_view_binary) as this
    package's callbacks, byte for byte, in 1D, 2D and 3D."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    lib = os.path.dirname(capi.LIBP4G)
    drv = os.path.join(ROOT, "tests", "boundary_driver.c")
    exes = {}
    for tag, hdir in (("ref", REFERENCE_CH6), ("prod", os.path.join(ROOT, "p4pdes_amd", "fish"))):
        flags = ["-std=c99", "-Wall", "-Werror", "-I" + inc, "-I" + hdir]
        _cc(flags + ["-c", os.path.join(hdir, "poissonfunctions.c"), "-o", "pf_%s.o" % tag], tmp_path)
        _cc(flags + ["-c", drv, "-o", "drv_%s.o" % tag], tmp_path)
        _cc(["drv_%s.o" % tag, "pf_%s.o" % tag, "-L" + lib, "-lp4g", "-Wl,-rpath," + lib, "-lm",
             "-o", "drv_%s" % tag], tmp_path)
        exes[tag] = str(tmp_path / ("drv_%s" % tag))
    for dim, m in ((1, 17), (2, 9), (3, 9)):
        outs = {}
        for tag, exe in exes.items():
            d = tmp_path / ("%s%d" % (tag, dim))
            d.mkdir()
            r = subprocess.run([exe, str(dim), str(m), str(d)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
            outs[tag] = ((d / "A.bin").read_bytes(), (d / "F.bin").read_bytes())
        assert outs["ref"][0] == outs["prod"][0], "operator differs in %dD" % dim
        assert outs["ref"][1] == outs["prod"][1], "residual differs in %dD" % dim
        F = capi.vec_load_binary(tmp_path / ("ref%d" % dim) / "F.bin")
        n = m * (m + 2 if dim > 1 else 1) * (m + 4 if dim > 2 else 1)
        assert F.size == n and np.count_nonzero(F) > n // 2
        # the AIJ header: class id, rows, columns, nonzeros (PETSc binary, big-endian int32)
        hdr = np.frombuffer(outs["ref"][0][:16], dtype=">i4")
        assert hdr[1] == n and hdr[2] == n and n < hdr[3] <= (2 * dim + 1) * n


def test_header_compiles_hosted_by_petsc(tmp_path):
    """With P4G_HOSTED_BY_PETSC (a program that also includes PETSc's own
    headers) p4g.h stands alone: the PETSc-shaped surface is left out and the
    solver ABI still compiles, as C and as C++."""
    src = tmp_path / "hosted.c"
    src.write_text("#define P4G_HOSTED_BY_PETSC 1\n#include \"p4g.h\"\n"
                   "#ifdef P4G_PETSC_H\n#error p4g_petsc.h was included\n#endif\n"
                   "int use(p4g_DA da, p4g_Solver s, const double *r, double *z) {\n"
                   "    p4g_Report rep; int n;\n"
                   "    if (p4g_solver_create(da, \"-pc_type mg\", &s)) return 1;\n"
                   "    if (p4g_pc_apply(s, r, z)) return 2;\n"
                   "    if (p4g_ksp_solve(s, r, z, &rep)) return 3;\n"
                   "    return p4g_device_count(&n);\n}\n")
    inc = "-I" + os.path.join(ROOT, "include")
    _cc(["-std=c99", "-Wall", "-Werror", inc, "-c", str(src), "-o", "h.o"], tmp_path)
    import subprocess
    r = subprocess.run(["g++", "-x", "c++", "-std=c++17", "-Wall", "-Werror", inc, "-c", str(src),
                        "-o", "hpp.o"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_mat_is_symmetric_checks_the_callback(tmp_path):
    """-mat_is_symmetric is a check of the entries the Jacobian callback
    inserts: the package's callbacks are symmetric; a copy with one coupling
    re-inserted 0.1% larger is not, at tolerances below that difference, and
    is within a tolerance above it."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    lib = os.path.dirname(capi.LIBP4G)
    fish = os.path.join(ROOT, "p4pdes_amd", "fish")
    exe = str(tmp_path / "drv")
    _cc(["-std=c99", "-Wall", "-Werror", "-I" + inc, "-I" + fish, os.path.join(ROOT, "tests", "boundary_driver.c"),
         os.path.join(fish, "poissonfunctions.c"), "-L" + lib, "-lp4g", "-Wl,-rpath," + lib, "-lm", "-o", exe],
        tmp_path)
    for dim, m in ((1, 9), (2, 9), (3, 5)):
        out = subprocess.run([exe, "sym", str(dim), str(m), "1e-7"], capture_output=True, text=True)
        assert out.returncode == 0, out.stderr
        assert out.stdout.split() == ["symmetric", "1", "skewed", "0"], (dim, out.stdout)
        out = subprocess.run([exe, "sym", str(dim), str(m), "1"], capture_output=True, text=True)
        assert out.stdout.split() == ["symmetric", "1", "skewed", "1"], (dim, out.stdout)
