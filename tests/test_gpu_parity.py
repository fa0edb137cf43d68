"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle
on the same options, grids and right-hand sides.

Bar (from the reference path's contract): the same CG iteration count
(within 1), the final solution within 1e-9 relative in the inf-norm, and
the same discretization error.  Full-size configurations are checked
through size-independent properties: the closed-form discretization error
of the sine eigen-modes, the true FP64 residual, and agreement of the
mixed-precision and FP64 device solutions.
"""
import math
import re

import numpy as np
import pytest

from oracle import oracle_py as O
from p4pdes_amd import capi

pytestmark = pytest.mark.gpu

MG = "-pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"

CASES = {
    "1d_manupoly_cheb2": "-fsh_dim 1 -fsh_problem manupoly -da_refine 5 %s -ksp_rtol 1e-12" % MG,
    "2d_manuexp_cheb1": "-fsh_dim 2 -da_refine 4 %s -mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG,
    "2d_c1_129": ("-fsh_dim 2 -fsh_problem manusin -da_refine 6 %s -pc_mg_levels 6 "
                  "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG),
    "3d_aniso_cheb2": ("-fsh_dim 3 -fsh_problem manupoly -fsh_cx 0.5 -fsh_cy 2 -fsh_cz 3 "
                       "-da_refine 4 %s -ksp_rtol 1e-10" % MG),
    "3d_sin_wcycle_richardson": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg "
                                 "-pc_mg_cycle_type w -mg_levels_ksp_type richardson "
                                 "-mg_levels_pc_type jacobi -mg_levels_ksp_max_it 1 "
                                 "-mg_levels_ksp_richardson_scale 0.8 -ksp_rtol 1e-10"),
    "3d_sin_cheb1_65": ("-fsh_dim 3 -fsh_problem manusin -da_refine 5 %s -pc_mg_levels 5 "
                        "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG),
    "3d_sin_cheb3": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -mg_levels_ksp_max_it 3 "
                     "-ksp_rtol 1e-10" % MG),
    "3d_sinemodes": ("-fsh_dim 3 -fsh_problem sinemodes -fsh_modes_seed 7 -da_refine 4 %s "
                     "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-8" % MG),
    "3d_noise_cheb2": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 4 %s "
                       "-ksp_rtol 1e-10" % MG),
    "3d_varcoef": ("-fsh_dim 3 -fsh_problem varcoef -da_refine 4 %s -mg_levels_ksp_max_it 1 "
                   "-ksp_rtol 1e-10" % MG),
    "3d_unprecond_norm": ("-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -ksp_norm_type "
                          "unpreconditioned -ksp_rtol 1e-10" % MG),
    "3d_natural_norm": ("-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -ksp_norm_type natural "
                        "-ksp_rtol 1e-10" % MG),
    "2d_no_gbdry_random": ("-fsh_dim 2 -fsh_initial_gonboundary false -fsh_initial_type random "
                           "-da_refine 3 %s -ksp_rtol 1e-10" % MG),
    "2d_pc_jacobi": "-fsh_dim 2 -da_refine 3 -pc_type jacobi -ksp_rtol 1e-10",
    "2d_pc_none": "-fsh_dim 2 -da_refine 2 -pc_type none -ksp_rtol 1e-10",
    "3d_one_level": "-fsh_dim 3 -fsh_problem manusin -da_refine 2 -pc_type mg -pc_mg_levels 1",
    "3d_cheb1_fused_cg": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -mg_levels_ksp_max_it 1 -ksp_rtol 1e-10 "
                          "-p4g_fuse_cg 1" % MG),
    "3d_cheb2_fused_cg_unprecond": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 4 %s "
                                    "-ksp_norm_type unpreconditioned -ksp_rtol 1e-10 -p4g_fuse_cg 1" % MG),
    "3d_richardson3_fused_cg": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -mg_levels_ksp_type "
                                "richardson -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 3 "
                                "-mg_levels_ksp_richardson_scale 0.7 -ksp_rtol 1e-10 -p4g_fuse_cg 1"),
    "3d_cheb1_fused_r_only": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -mg_levels_ksp_max_it 1 "
                              "-ksp_rtol 1e-10 -p4g_fuse_r 1 -p4g_fuse_dir 0" % MG),
    "3d_cheb2_fused_dir_only": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 4 %s "
                                "-ksp_rtol 1e-10 -p4g_fuse_r 0 -p4g_fuse_dir 1" % MG),
    "3d_galerkin_cheb2": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_galerkin "
                          "-ksp_rtol 1e-10" % MG),
    "3d_galerkin_cheb1_unprecond": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 4 %s "
                                    "-pc_mg_galerkin -mg_levels_ksp_max_it 1 -ksp_norm_type unpreconditioned "
                                    "-ksp_rtol 1e-10" % MG),
    "3d_galerkin_test7_problem_jacobi": ("-fsh_dim 3 -fsh_problem manupoly -ksp_rtol 1.0e-12 -pc_type mg "
                                         "-pc_mg_galerkin -da_refine 2 -mg_levels_pc_type jacobi"),
    "2d_galerkin_w_richardson": ("-fsh_dim 2 -fsh_problem manupoly -fsh_cx 0.3 -da_refine 5 -pc_type mg "
                                 "-pc_mg_galerkin both -pc_mg_cycle_type w -mg_levels_ksp_type richardson "
                                 "-mg_levels_pc_type jacobi -mg_levels_ksp_richardson_scale 0.6 "
                                 "-mg_levels_ksp_max_it 2 -ksp_rtol 1e-10"),
    "1d_galerkin": "-fsh_dim 1 -fsh_problem manupoly -da_refine 6 %s -pc_mg_galerkin -ksp_rtol 1e-10" % MG,
    "3d_fcycle_cheb1": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_type full "
                        "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG),
    "2d_fcycle_w_cheb2": ("-fsh_dim 2 -fsh_problem manupoly -da_refine 5 %s -pc_mg_type full "
                          "-pc_mg_cycle_type w -ksp_rtol 1e-10" % MG),
    "3d_fcycle_galerkin": ("-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -pc_mg_type full -pc_mg_galerkin "
                           "-ksp_rtol 1e-10" % MG),
    "1d_fcycle": "-fsh_dim 1 -fsh_problem manupoly -da_refine 6 %s -pc_mg_type full -ksp_rtol 1e-10" % MG,
    "3d_richardson3": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -mg_levels_ksp_type "
                       "richardson -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 3 "
                       "-mg_levels_ksp_richardson_scale 0.7 -ksp_rtol 1e-10"),
    # outer Richardson: the multigrid shortcut (no monitor) and PETSc's
    # generic loop (with -ksp_monitor), cycle variants of c/ch6/study
    "3d_outer_richardson_mg_shortcut": "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -ksp_type richardson "
                                       "-ksp_rtol 1e-10" % MG,
    "2d_outer_richardson_unprecond_w": ("-fsh_dim 2 -fsh_problem manupoly -da_refine 5 %s -ksp_type richardson "
                                        "-ksp_norm_type unpreconditioned -ksp_monitor -pc_mg_cycle_type w "
                                        "-ksp_rtol 1e-10" % MG),
    "3d_outer_richardson_scaled_jacobi": ("-fsh_dim 3 -fsh_problem manusin -da_refine 3 -pc_type jacobi "
                                          "-ksp_type richardson -ksp_richardson_scale 0.9 -ksp_rtol 1e-3 "
                                          "-ksp_max_it 20000"),
    "3d_mcycles2_cg": "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_multiplicative_cycles 2 "
                      "-ksp_rtol 1e-10" % MG,
    "3d_kaskade_cg": "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_type kaskade "
                     "-mg_levels_ksp_max_it 4 -ksp_rtol 1e-10" % MG,
    # the LDS-staged 3D passes (k_stencil_lds) forced on where rows fill the
    # waves (FP64: 129^3), with and without the fused CG updates
    "3d_lds_cheb2_129": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 6 %s "
                         "-ksp_norm_type unpreconditioned -ksp_rtol 1e-10 -p4g_lds 1" % MG),
    "3d_lds_cheb1_fused_cg_129": ("-fsh_dim 3 -fsh_problem manusin -da_refine 6 %s -mg_levels_ksp_max_it 1 "
                                  "-ksp_rtol 1e-10 -p4g_fuse_cg 1 -p4g_lds 1" % MG),
    "3d_kaskade_outer_richardson": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_type kaskade "
                                    "-mg_levels_ksp_max_it 4 -ksp_type richardson -ksp_monitor -ksp_rtol 1e-10" % MG),
}


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", sorted(CASES))
def test_fish_parity(name):
    args = CASES[name]
    print("case", name, args, flush=True)
    rc, text, res, u = capi.fish_run(args, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(args, want_u=True)
    assert orc == 0, otext
    assert res.reason > 0, text
    assert abs(res.its - ores.its) <= 1, (res.its, ores.its)
    assert _rel(u, ou) < 1e-9
    # |e_gpu - e_oracle|_inf <= |u_gpu - u_oracle|_inf (triangle inequality)
    assert abs(res.errinf - ores.errinf) <= float(np.max(np.abs(u - ou))) * (1 + 1e-9) + 1e-15
    # printed error line: same text as the oracle's (%.3e)
    assert text.splitlines()[-1] == otext.splitlines()[-1]


# V-cycles that are indefinite preconditioners (Richardson smoothing at a
# negative or too large scale): CG stops on the sign change of beta = (z, r)
# with DIVERGED_INDEFINITE_PC, at the oracle's iteration and with its iterate
INDEFINITE_PC = {
    "2d_richardson_negative_scale": ("-fsh_dim 2 -da_refine 3 -pc_type mg -mg_levels_ksp_type richardson "
                                     "-mg_levels_ksp_richardson_scale -1 -mg_levels_pc_type jacobi"),
    "3d_richardson_scale_2.5": ("-fsh_dim 3 -da_refine 3 -pc_type mg -mg_levels_ksp_type richardson "
                                "-mg_levels_ksp_richardson_scale 2.5 -mg_levels_pc_type jacobi"),
}


@pytest.mark.parametrize("variant", ["", " -p4g_x_defer 0"])
@pytest.mark.parametrize("name", sorted(INDEFINITE_PC))
def test_indefinite_preconditioner_stops_as_oracle(name, variant):
    args = INDEFINITE_PC[name] + " -ksp_converged_reason"
    print("case", name, variant, flush=True)
    rc, text, res, u = capi.fish_run(args + variant, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(args, want_u=True)
    assert orc == 0, otext
    assert ores.reason == -8 and res.reason == -8, (res.reason, ores.reason)
    assert res.its == ores.its
    assert _rel(u, ou) < 1e-8
    assert [l for l in text.splitlines() if "Linear solve" in l] == \
        [l for l in otext.splitlines() if "Linear solve" in l]


def test_mixed_indefinite_preconditioner_stops():
    rc, text, res, _ = capi.fish_run(INDEFINITE_PC["3d_richardson_scale_2.5"] + " -p4g_precision mixed")
    assert rc == 0, text
    assert res.reason == -8, res.reason


@pytest.mark.parametrize("name", sorted(n for n in CASES if "galerkin" in n))
def test_galerkin_trajectory(name):
    # coarse operators P^T A P formed on the device in 27-point form against
    # the oracle's sparse triple product: the same iteration count and the
    # same final residual reduction (the coarse operators steer the whole
    # residual history, not only the count)
    args = CASES[name]
    rc, text, res, _ = capi.fish_run(args)
    assert rc == 0, text
    orc, otext, ores, _ = O.fish_run(args)
    assert orc == 0, otext
    assert res.its == ores.its, (res.its, ores.its)
    red, ored = res.rnorm / res.rnorm0, ores.rnorm / ores.rnorm0
    print(name, "its", res.its, "reduction", red, "oracle", ored, flush=True)
    assert abs(red - ored) <= 1e-8 * ored


def test_galerkin_off_is_rediscretization():
    # -pc_mg_galerkin none leaves the rediscretized coarse operators
    base = "-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -ksp_rtol 1e-10" % MG
    rc0, t0, r0, u0 = capi.fish_run(base, want_u=True)
    rc1, t1, r1, u1 = capi.fish_run(base + " -pc_mg_galerkin none", want_u=True)
    assert rc0 == 0 and rc1 == 0
    assert r0.its == r1.its and np.array_equal(u0, u1)


def test_golden_lines_with_jacobi_smoother():
    # the golden fish.test1 problem with a Jacobi-Chebyshev smoother: the
    # discretization error line matches the reference's exactly
    args = ("-fsh_dim 1 -fsh_problem manupoly -da_refine 3 -pc_type mg -mg_levels_pc_type jacobi "
            "-ksp_rtol 1.0e-12 -snes_monitor_short -ksp_converged_reason")
    rc, text, res, _ = capi.fish_run(args)
    assert rc == 0, text
    lines = text.splitlines()
    assert lines[0] == "  0 SNES Function norm 0.925546"
    assert lines[2] == "  1 SNES Function norm < 1.e-11"
    assert lines[3:] == ["problem manupoly on 17 point 1D grid:",
                         "  error |u-uexact|_inf = 9.766e-04, |u-uexact|_h = 7.132e-04"]
    orc, otext, ores, _ = O.fish_run(args)
    assert otext == text


def _golden_fish():
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return {t["test"]: t for t in json.load(open(os.path.join(here, "golden", "fish_golden.json")))["tests"]}


@pytest.mark.parametrize("num", [1, 2, 3, 4, 5, 6, 7, 8])
def test_golden_fish_on_device(num):
    # the reference's own commands (PETSc's defaults: ILU(0) as the KSP's PC,
    # Chebyshev/SOR level smoothing), byte for byte against
    # c/ch6/output/fish.testN; the 2-rank tests with the local SOR sweeps of
    # two processes
    t = _golden_fish()[num]
    args = t["options"] + (" -p4g_sor_blocks %d" % t["nprocs"] if t["nprocs"] > 1 else "")
    print("golden fish", num, args, flush=True)
    rc, text, res, _ = capi.fish_run(args)
    assert rc == 0, text
    assert text == t["expected"]


SOR_CASES = {
    "3d_cheb_sor": "-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -ksp_rtol 1e-10",
    "2d_w_rich_sor_2blocks": ("-fsh_dim 2 -fsh_problem manupoly -da_refine 5 -pc_type mg -pc_mg_cycle_type w "
                              "-mg_levels_ksp_type richardson -mg_levels_ksp_max_it 1 -ksp_rtol 1e-10 "
                              "-p4g_sor_blocks 2 -oracle_nranks 2"),
    "3d_galerkin_sor_3blocks": ("-fsh_dim 3 -fsh_problem manupoly -da_refine 3 -pc_type mg -pc_mg_galerkin "
                                "-ksp_rtol 1e-10 -p4g_sor_blocks 3 -oracle_nranks 3"),
    "2d_fcycle_sor": "-fsh_dim 2 -fsh_problem manusin -da_refine 5 -pc_type mg -pc_mg_type full -ksp_rtol 1e-10",
    "3d_kaskade_richardson_sor": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -pc_mg_type kaskade "
                                  "-mg_levels_ksp_max_it 4 -ksp_type richardson -ksp_monitor -ksp_rtol 1e-8"),
    "1d_sor": "-fsh_dim 1 -fsh_problem manupoly -da_refine 7 -pc_type mg -ksp_rtol 1e-12",
}


SOR_CASES.update({
    "3d_cg_ilu": "-fsh_dim 3 -fsh_problem manusin -da_refine 4 -ksp_rtol 1e-10",
    "2d_cg_icc_aniso": "-fsh_dim 2 -fsh_problem manupoly -fsh_cx 100 -da_refine 6 -pc_type icc -ksp_rtol 1e-10",
    "2d_mg_cheb_icc_aniso": ("-fsh_dim 2 -fsh_problem manupoly -fsh_cy 1000 -da_refine 6 -pc_type mg "
                             "-mg_levels_pc_type icc -ksp_rtol 1e-10"),
    "3d_mg_rich_ilu": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -mg_levels_ksp_type richardson "
                       "-mg_levels_pc_type ilu -mg_levels_ksp_max_it 1 -ksp_rtol 1e-10"),
})


# -ksp_type gmres: restarted GMRES(m), left preconditioning (PETSc's default
# KSP; the oracle's restatement is pinned in tests/test_minimal.py)
SOR_CASES.update({
    "2d_gmres_mg_jacobi": ("-fsh_dim 2 -fsh_problem manupoly -da_refine 5 -ksp_type gmres -pc_type mg "
                           "-mg_levels_pc_type jacobi -ksp_rtol 1e-10"),
    "3d_gmres_ilu_restart5": "-fsh_dim 3 -fsh_problem manusin -da_refine 3 -ksp_type gmres -ksp_gmres_restart 5 "
                             "-ksp_rtol 1e-10",
    "2d_gmres_none_restarts": "-fsh_dim 2 -fsh_problem manupoly -fsh_cx 10 -da_refine 4 -ksp_type gmres -pc_type none "
                              "-ksp_rtol 1e-8",
})


# c/ch6/study/coarseperformance.sh: a fixed count of coarse PCG iterations
# (-mg_coarse_ksp_type cg -mg_coarse_ksp_convergence_test skip) in place of
# the coarse LU; the W-cycle restarts the coarse CG from a nonzero iterate
_CCG = "-mg_coarse_ksp_type cg -mg_coarse_ksp_convergence_test skip"
SOR_CASES.update({
    "3d_coarse_cg_icc": ("-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -pc_mg_levels 3 %s "
                         "-mg_coarse_pc_type icc -mg_coarse_ksp_max_it 2 -ksp_rtol 1e-10" % _CCG),
    "2d_coarse_cg_jacobi_w": ("-fsh_dim 2 -fsh_problem manupoly -da_refine 6 -pc_type mg -pc_mg_levels 3 "
                              "-pc_mg_cycle_type w -mg_levels_pc_type jacobi %s -mg_coarse_pc_type jacobi "
                              "-mg_coarse_ksp_max_it 5 -ksp_rtol 1e-10" % _CCG),
    "1d_coarse_cg_none": ("-fsh_dim 1 -fsh_problem manupoly -da_refine 8 -pc_type mg -pc_mg_levels 4 %s "
                          "-mg_coarse_pc_type none -mg_coarse_ksp_max_it 3 -ksp_rtol 1e-12" % _CCG),
    "3d_coarse_cg_fcycle": ("-fsh_dim 3 -fsh_problem manupoly -da_refine 4 -pc_type mg -pc_mg_type full "
                            "-pc_mg_levels 3 -mg_levels_pc_type jacobi %s -mg_coarse_pc_type ilu "
                            "-mg_coarse_ksp_max_it 3 -ksp_rtol 1e-10" % _CCG),
})


@pytest.mark.parametrize("name", sorted(SOR_CASES))
def test_sor_smoother_parity(name):
    args = SOR_CASES[name]
    # the device reads -p4g_sor_blocks, the oracle -oracle_nranks
    dev = re.sub(r"-oracle_nranks \d+", "", args)
    print("sor case", name, dev, flush=True)
    rc, text, res, u = capi.fish_run(dev, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(args, want_u=True)
    assert orc == 0, otext
    assert res.reason > 0 and abs(res.its - ores.its) <= 1, (res.its, ores.its)
    assert _rel(u, ou) < 1e-9
    assert text.splitlines()[-1] == otext.splitlines()[-1]


@pytest.mark.parametrize("depth", [4, 3])
def test_coarse_cg_depth_study(depth):
    # c/ch6/study/coarseperformance.sh at 129^3: two ICC-preconditioned CG
    # iterations on the coarsest grid (33^3 at depth 3) raise the V-cycle
    # count when the cycle is shallow
    args = ("-fsh_dim 3 -ksp_rtol 1.0e-10 -ksp_converged_reason -pc_type mg -da_refine 6 -pc_mg_levels %d "
            "-mg_coarse_ksp_type cg -mg_coarse_pc_type icc -mg_coarse_ksp_max_it 2 "
            "-mg_coarse_ksp_convergence_test skip" % depth)
    print("coarse study", depth, flush=True)
    rc, text, res, u = capi.fish_run(args, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(args, want_u=True)
    assert orc == 0, otext
    assert res.reason == 2 and abs(res.its - ores.its) <= 1, (res.its, ores.its)
    assert _rel(u, ou) < 1e-9
    assert text.splitlines()[-1] == otext.splitlines()[-1]


OPS = [
    "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -ksp_rtol 1e-10" % MG,
    "-fsh_dim 2 -fsh_problem manupoly -fsh_cx 0.3 -da_refine 5 %s -mg_levels_ksp_max_it 1" % MG,
    "-fsh_dim 3 -fsh_problem varcoef -da_refine 3 %s -mg_levels_ksp_max_it 1" % MG,
    "-fsh_dim 3 -fsh_problem manusin -da_refine 3 -pc_type mg -pc_mg_cycle_type w "
    "-mg_levels_ksp_type richardson -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 2",
]


@pytest.mark.parametrize("args", OPS, ids=["3d", "2d_aniso", "3d_varcoef", "3d_w_rich"])
def test_operator_and_vcycle_parity(args):
    print("case", args, flush=True)
    with capi.Session(args) as s:
        sv = capi.Solver(s)
        rng = np.random.default_rng(5)
        x = rng.standard_normal(s.n)
        y = sv.mat_mult(x)
        assert _rel(y, O.mat_mult(args, x)) < 1e-14
        z = sv.pc_apply(x)
        assert _rel(z, O.pc_apply(args, x)) < 1e-11
        # the V-cycle operator is symmetric: <B x, w> = <x, B w>
        w = rng.standard_normal(s.n)
        bw = sv.pc_apply(w)
        assert abs(z @ w - x @ bw) < 1e-11 * abs(z @ w)


@pytest.mark.parametrize("args", [
    "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_type full -mg_levels_ksp_max_it 1" % MG,
    "-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -pc_mg_galerkin" % MG,
    "-fsh_dim 2 -fsh_problem manupoly -da_refine 5 %s -pc_mg_galerkin -pc_mg_type full "
    "-pc_mg_cycle_type w" % MG,
    "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -pc_mg_type kaskade -mg_levels_ksp_max_it 3" % MG,
    "-fsh_dim 2 -fsh_problem manupoly -da_refine 5 %s -pc_mg_multiplicative_cycles 3 -pc_mg_cycle_type w" % MG],
    ids=["3d_fcycle", "3d_galerkin", "2d_galerkin_fcycle_w", "3d_kaskade", "2d_mcycles3_w"])
def test_fcycle_and_galerkin_apply_parity(args):
    # one preconditioner application against the oracle's, and for the
    # (symmetric) Galerkin V-cycle the symmetry of B
    print("case", args, flush=True)
    with capi.Session(args) as s:
        sv = capi.Solver(s)
        rng = np.random.default_rng(7)
        x = rng.standard_normal(s.n)
        z = sv.pc_apply(x)
        assert _rel(z, O.pc_apply(args, x)) < 1e-11
        if "full" not in args and "kaskade" not in args:
            w = rng.standard_normal(s.n)
            assert abs(z @ w - x @ sv.pc_apply(w)) < 1e-11 * abs(z @ w)


def test_ksp_viewer_files(tmp_path):
    # -ksp_view_mat/-ksp_view_rhs/-ksp_view_solution binary:<file> after the
    # solve: the operator and F(u0) equal the oracle's, and the solution
    # solves the system to the requested tolerance (PETSc binary layout)
    from test_binary_io import read_petsc, _csr_apply
    fA, fb = tmp_path / "A.dat", tmp_path / "by.dat"
    args = ("-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -ksp_rtol 1e-10 -ksp_norm_type unpreconditioned "
            "-ksp_view_mat binary:%s -ksp_view_rhs binary:%s -ksp_view_solution binary:%s" % (MG, fA, fb, fb))
    rc, text, res, _ = capi.fish_run(args)
    assert rc == 0, text
    (_, (M, N, rowptr, cols, vals)), = read_petsc(fA)
    (k1, b), (k2, y) = read_petsc(fb)
    assert k1 == k2 == "vec" and b.size == y.size == M
    plain = " ".join(a for a in args.split() if not a.startswith("binary:") and not a.startswith("-ksp_view"))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(M)
    ax = _csr_apply(M, rowptr, cols, vals, x)
    assert np.max(np.abs(ax - O.mat_mult(plain, x))) <= 1e-12 * np.max(np.abs(ax))
    assert _rel(b, O.rhs(plain)) < 1e-14
    r = b - _csr_apply(M, rowptr, cols, vals, y)
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b)   # recursive vs true residual
    assert np.array_equal(capi.vec_load_binary(fb, 1), y)


@pytest.mark.parametrize("args", [
    "-fsh_dim 3 -fsh_problem manusin -da_refine 4 %s -ksp_rtol 1e-10" % MG,
    "-fsh_dim 2 -fsh_problem manupoly -da_refine 5 %s -ksp_norm_type unpreconditioned" % MG,
    "-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s -ksp_norm_type natural -pc_mg_galerkin" % MG,
    # c/ch6/study/cyclereductions.sh: one Richardson iteration per cycle type
    "-fsh_dim 3 -da_refine 4 %s -ksp_type richardson -ksp_max_it 1 -ksp_norm_type unpreconditioned "
    "-mg_levels_ksp_max_it 2 -pc_mg_cycle_type w" % MG,
    "-fsh_dim 3 -da_refine 4 %s -ksp_type richardson -ksp_max_it 1 -ksp_norm_type unpreconditioned "
    "-mg_levels_ksp_max_it 1 -pc_mg_multiplicative_cycles 2 -fsh_initial_type random" % MG,
    "-fsh_dim 3 -da_refine 4 %s -ksp_type richardson -ksp_max_it 1 -ksp_norm_type unpreconditioned "
    "-mg_levels_ksp_max_it 2 -pc_mg_type full" % MG,
    "-fsh_dim 3 -da_refine 4 %s -ksp_type richardson -ksp_max_it 1 -ksp_norm_type unpreconditioned "
    "-mg_levels_ksp_max_it 4 -pc_mg_type kaskade" % MG,
    # GMRES: the Hessenberg estimates at PETSc's monitor points, restarts included
    "-fsh_dim 3 -fsh_problem manusin -da_refine 3 -ksp_type gmres -ksp_gmres_restart 5 -ksp_rtol 1e-10",
    "-fsh_dim 2 -fsh_problem manupoly -da_refine 5 -ksp_type gmres %s -ksp_rtol 1e-10" % MG,
], ids=["3d_precond", "2d_unprecond", "3d_natural_galerkin", "study_w", "study_mcycles2_random",
        "study_full", "study_kaskade", "gmres_ilu_restart5", "gmres_mg"])
def test_ksp_monitor_lines(args):
    # -ksp_monitor: PETSc's KSPMonitorResidual lines, before the converged
    # reason, with the oracle's residual norms
    a = args + " -ksp_monitor -ksp_converged_reason -snes_monitor_short"
    rc, text, res, _ = capi.fish_run(a)
    assert rc == 0, text
    orc, otext, ores, _ = O.fish_run(a)
    g = [l for l in text.splitlines() if "KSP Residual norm" in l]
    o = [l for l in otext.splitlines() if "KSP Residual norm" in l]
    assert len(g) == len(o) == res.its + 1, (len(g), len(o), res.its)
    pat = re.compile(r"^     *\d+ KSP Residual norm \d\.\d{12}e[+-]\d\d $")
    h0 = float(o[0].split()[-1])
    for q, (lg, lo) in enumerate(zip(g, o)):
        assert pat.match(lg), lg
        assert lg.split("KSP")[0] == lo.split("KSP")[0]
        vg, vo = float(lg.split()[-1]), float(lo.split()[-1])
        assert abs(vg - vo) <= 1e-6 * abs(vo) + 1e-12 * h0, (q, vg, vo)
    lines = text.splitlines()
    assert lines.index(g[-1]) + 1 == next(i for i, l in enumerate(lines) if "Linear solve" in l)


def test_rhs_parity():
    args = "-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 3 %s" % MG
    with capi.Session(args) as s:
        F = np.zeros(s.n)
        capi.check(capi.lib().p4g_memcpy_dtoh(F.ctypes.data, capi.fishlib().fish_session_rhs_device(s.h),
                                              8 * s.n))
    assert _rel(F, O.rhs(args)) < 1e-14


def _eig_err_3d(m):
    h = 1.0 / (m - 1)
    return abs(3 * math.pi ** 2 / ((12 / h ** 2) * math.sin(math.pi * h / 2) ** 2) - 1)


def test_c2_257_fp64_closed_form():
    # C2: 257^3, 7 levels to 5^3 Cholesky, FP64, rtol 1e-10
    args = ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -pc_mg_levels 7 "
            "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG)
    with capi.Session(args) as s:
        rep = s.solve()
        print("C2 its", rep.its, "solve_ms", rep.solve_ms, "relres", rep.true_relres, flush=True)
        assert rep.reason == 2
        einf, _ = s.errors()
    assert abs(einf - 1.254995e-05) < 5e-4 * 1.254995e-05
    assert abs(_eig_err_3d(257) - 1.254995e-05) < 1e-11
    # the test is on the preconditioned norm (PETSc's CG default); for this
    # smooth right-hand side the true residual falls less (B ~ A^-1 shrinks
    # the rough residual left at the end relative to the smooth initial one)
    assert rep.true_relres < 1e-6
    assert rep.rnorm <= 1e-10 * rep.rnorm0


@pytest.mark.parametrize("refine", [4, 5])
def test_mixed_matches_fp64(refine):
    base = ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine %d %s -pc_mg_levels %d "
            "-mg_levels_ksp_max_it 2 -ksp_rtol 1e-10" % (refine, MG, refine))
    rc, text, res, u = capi.fish_run(base + " -p4g_precision mixed", want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(base + " -ksp_norm_type unpreconditioned", want_u=True)
    assert orc == 0
    assert res.true_relres <= 1e-10 * 1.0001
    # same discretization error (4 significant digits) and a small distance
    # to the FP64 oracle solution
    assert abs(res.errinf - ores.errinf) < 1e-4 * ores.errinf
    assert _rel(u, ou) < 1e-7
    print("mixed: outer its", res.its, "inner its", res.inner_its, "rel diff to FP64 oracle",
          _rel(u, ou))


@pytest.mark.parametrize("variant", ["-p4g_mixed_x64 0", "-p4g_mixed_x64 1", "-p4g_mixed_x64 0 -p4g_fuse_cg 1",
                                     "-p4g_mixed_x64 1 -p4g_fuse_cg 1", "-p4g_x_defer 0", "-p4g_x_defer 1",
                                     "-p4g_x_defer 15 -p4g_reliable_delta 0.001",
                                     "-p4g_fp32_diff_form 0", "-p4g_reliable_delta 0.01",
                                     "-fsh_problem varcoef", "-p4g_fuse_cg 1", "-p4g_fuse_cg 1 -fsh_problem varcoef",
                                     "-p4g_fuse_r 1 -p4g_fuse_dir 0", "-p4g_fuse_r 0 -p4g_fuse_dir 1",
                                     "-pc_mg_type full"])
def test_mixed_variants(variant):
    # every mixed-precision variant converges on the FP64 true residual to the
    # same discrete solution as the FP64 oracle
    base = ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 4 %s -pc_mg_levels 4 "
            "-mg_levels_ksp_max_it 2 -ksp_rtol 1e-10 %s" % (MG, variant))
    print("case", variant, flush=True)
    rc, text, res, u = capi.fish_run(base + " -p4g_precision mixed", want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(base + " -ksp_norm_type unpreconditioned", want_u=True)
    assert orc == 0
    assert res.reason > 0
    assert res.true_relres <= 1e-10 * 1.0001
    assert abs(res.errinf - ores.errinf) < 1e-4 * ores.errinf
    assert _rel(u, ou) < 1e-7
    print("  its", res.its, "oracle its", ores.its, "replacements", res.inner_its, flush=True)


@pytest.mark.parametrize("case", ["-da_refine 5 -pc_mg_levels 4", "-da_refine 7 -p4g_fuse_cg 1",
                                  "-da_refine 6 -p4g_reliable_delta 0.001 -pc_mg_levels 5"])
def test_x_defer_matches_per_iteration_updates(case):
    # the FP64 iterate updates deferred to the replacement pass add the same
    # alpha_j p_j in the same order as the per-iteration updates of the
    # direction pass: the same solve bit for bit
    base = ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 %s %s -ksp_rtol 1e-10 "
            "-p4g_precision mixed" % (MG, case))
    print("case", case, flush=True)
    rc0, t0, r0, u0 = capi.fish_run(base + " -p4g_x_defer 0", want_u=True)
    rc1, t1, r1, u1 = capi.fish_run(base, want_u=True)
    assert rc0 == 0 and rc1 == 0, (t0, t1)
    assert r0.reason > 0 and r0.its == r1.its and r0.inner_its == r1.inner_its
    print("  its", r1.its, "replacements", r1.inner_its, flush=True)
    assert np.array_equal(u0, u1), _rel(u1, u0)


@pytest.mark.parametrize("case", ["-fsh_dim 3 -da_refine 5 -pc_type mg -pc_mg_levels 4 -mg_levels_ksp_type chebyshev "
                                  "-mg_levels_pc_type jacobi -ksp_rtol 1e-10",
                                  "-fsh_dim 3 -da_refine 7 -pc_type mg -mg_levels_ksp_type chebyshev "
                                  "-mg_levels_pc_type jacobi -ksp_rtol 1e-10 -p4g_fuse_cg 1",
                                  "-fsh_dim 2 -da_refine 6 -pc_type jacobi -ksp_rtol 1e-8",
                                  "-fsh_dim 2 -da_refine 5 -pc_type none -ksp_norm_type unpreconditioned "
                                  "-ksp_rtol 1e-8"])
def test_fp64_x_defer_matches_per_iteration_updates(case):
    # FP64 CG with the iterate updates deferred (up to 7 listed, applied in
    # iteration order when the list is full and on exit) adds the same
    # alpha_j p_j in the same order as the direction pass would: the same
    # solve bit for bit (the Jacobi and unpreconditioned cases fold many times)
    print("case", case, flush=True)
    rc0, t0, r0, u0 = capi.fish_run("-fsh_problem manusin " + case + " -p4g_x_defer 0", want_u=True)
    rc1, t1, r1, u1 = capi.fish_run("-fsh_problem manusin " + case, want_u=True)
    assert rc0 == 0 and rc1 == 0, (t0, t1)
    assert r0.its == r1.its and r0.reason == r1.reason
    print("  its", r1.its, "reason", r1.reason, flush=True)
    assert np.array_equal(u0, u1), _rel(u1, u0)
    assert t0 == t1


def test_c4_varcoef_257_mixed_and_fp64():
    # C4 at full size: both precisions reach the FP64 true residual, agree with
    # each other, and the discretization error falls ~4x from 129^3 (O(h^2))
    base = ("-fsh_dim 3 -fsh_problem varcoef -da_refine %d -pc_mg_levels %d %s -mg_levels_ksp_max_it 1 "
            "-ksp_rtol 1e-10")
    res = {}
    for r in (6, 7):
        for prec in ("-ksp_norm_type unpreconditioned", "-p4g_precision mixed"):
            rc, text, fr, u = capi.fish_run(base % (r, r, MG) + " " + prec, want_u=True)
            assert rc == 0, text
            assert fr.reason > 0 and fr.true_relres <= 1e-10 * 1.0001, (r, prec, fr.true_relres)
            res[(r, prec)] = (fr, u)
            print("refine", r, prec, "its", fr.its, "errinf", fr.errinf, flush=True)
    for r in (6, 7):
        (f64, u64), (fm, um) = res[(r, "-ksp_norm_type unpreconditioned")], res[(r, "-p4g_precision mixed")]
        assert _rel(um, u64) < 1e-7
        assert abs(fm.errinf - f64.errinf) < 1e-4 * f64.errinf
    ratio = res[(6, "-ksp_norm_type unpreconditioned")][0].errinf / res[(7, "-ksp_norm_type unpreconditioned")][0].errinf
    assert 3.0 < ratio < 5.0, ratio


LDS_PAIRS = {
    "fp64_129_cheb2_fused": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 6 %s "
                             "-ksp_norm_type unpreconditioned -ksp_rtol 1e-10 -p4g_fuse_cg 1" % MG),
    "fp64_129_cheb1": ("-fsh_dim 3 -fsh_problem manusin -da_refine 6 %s -mg_levels_ksp_max_it 1 "
                       "-ksp_rtol 1e-10" % MG),
    "fp64_257_cheb1": ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -mg_levels_ksp_max_it 1 "
                       "-ksp_rtol 1e-10" % MG),
    "mixed_257_cheb2": ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -da_refine 7 %s "
                        "-ksp_rtol 1e-10 -p4g_precision mixed" % MG),
    "mixed_257_cheb1_fused": ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -mg_levels_ksp_max_it 1 "
                              "-ksp_rtol 1e-10 -p4g_precision mixed -p4g_fuse_cg 1" % MG),
    "mixed_257_cheb2_fused_x32": ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -ksp_rtol 1e-10 "
                                  "-p4g_precision mixed -p4g_fuse_cg 1 -p4g_mixed_x64 0" % MG),
    "mixed_257_cheb2_fused_x64": ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -ksp_rtol 1e-10 "
                                  "-p4g_precision mixed -p4g_fuse_cg 1 -p4g_mixed_x64 1" % MG),
}


@pytest.mark.parametrize("name", sorted(LDS_PAIRS))
def test_lds_passes_match_register_passes(name):
    # k_stencil_lds forms the same operand with the same operation order as
    # k_stencil, and its blocks sum their dot products in the same order:
    # with two smoothing steps (separate prolongation) the solve is the same
    # bit for bit with the LDS staging on or off.  With one step the fused
    # prolongation + smoothing runs k_prolong_smooth_lds, whose blocks (8
    # output rows, not 6) sum the cycle's dot products in another order:
    # the same iterations, the solutions within rounding.
    args = LDS_PAIRS[name]
    print("lds pair", name, args, flush=True)
    rc0, t0, r0, u0 = capi.fish_run(args + " -p4g_lds 0", want_u=True)
    rc1, t1, r1, u1 = capi.fish_run(args + " -p4g_lds 1", want_u=True)
    assert rc0 == 0 and rc1 == 0, (t0, t1)
    assert r0.its == r1.its and r0.reason > 0
    if ("cheb1" in name and "fp64" in name) or "mixed" in name:
        # FP32 passes: the fused residual + restriction computes the same
        # per-point formulas with two points per lane where the separate
        # FP32 passes take four, so the compiler's FMA contractions may
        # round differently; FP64 V(1,1): another dot-product order
        assert _rel(u1, u0) < 1e-12
        assert t0.splitlines()[-1] == t1.splitlines()[-1]
    else:
        assert np.array_equal(u0, u1), float(np.max(np.abs(u0 - u1)))


@pytest.mark.parametrize("prec", ["double", "mixed"])
def test_fused_resid_restrict_pc_apply(prec):
    # one multigrid cycle z = B r with the residual + restriction fused
    # (k_resid_restrict_lds) and as two passes: the same formulas, whose FMA
    # contractions the compiler chooses per kernel, so the two agree to
    # rounding (FP64: 1.7e-16 relative measured at 257^3; FP32: its own eps)
    args = ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -mg_levels_ksp_max_it 2 -p4g_lds 1 "
            "-p4g_precision %s" % (MG, prec))
    r = np.random.default_rng(5).standard_normal(257 ** 3)
    zs = []
    for mode in ("0", "1"):
        print("resid_restrict", prec, mode, flush=True)
        with capi.Session(args + " -p4g_resid_restrict " + mode) as sess:
            zs.append(capi.Solver(sess).pc_apply(r))
    assert _rel(zs[1], zs[0]) < (1e-14 if prec == "double" else 1e-5)


@pytest.mark.parametrize("prec", ["double", "mixed"])
def test_fused_prolong_smooth_pc_apply(prec):
    # one multigrid cycle z = B r with the prolongation fused into the first
    # of two Chebyshev post-smoothing steps (k_prolong_smooth_lds, two fine
    # points per lane, the prolonged iterate kept for the second step) and as
    # separate passes: the same per-point formulas up to the FMA
    # contractions the compiler picks per kernel
    args = ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 %s -mg_levels_ksp_max_it 2 -p4g_lds 1 "
            "-p4g_precision %s" % (MG, prec))
    r = np.random.default_rng(6).standard_normal(257 ** 3)
    zs = []
    for mode in ("0", "1"):
        print("post_fuse", prec, mode, flush=True)
        with capi.Session(args + " -p4g_post_fuse " + mode) as sess:
            zs.append(capi.Solver(sess).pc_apply(r))
    assert _rel(zs[1], zs[0]) < (1e-14 if prec == "double" else 1e-5)


SOR3_PAIRS = {
    "65_cheb_sor": "-fsh_dim 3 -fsh_problem manusin -da_refine 5 -pc_type mg -ksp_rtol 1e-10",
    "129_cheb_sor_w": "-fsh_dim 3 -fsh_problem manupoly -da_refine 6 -pc_type mg -pc_mg_cycle_type w -ksp_rtol 1e-8",
}


@pytest.mark.parametrize("name", sorted(SOR3_PAIRS))
def test_sor3_pipelined_matches_hyperplanes(name):
    # PETSc's default level smoother on 3D grids: the pipelined strip sweeps
    # (k_sor7_strips) update every point from the same values in the same
    # order as the hyperplane sweeps (k_sor_plane): the same solve bit for bit
    args = SOR3_PAIRS[name]
    print("sor3 pair", name, args, flush=True)
    rc0, t0, r0, u0 = capi.fish_run(args + " -p4g_sor_pipeline 0", want_u=True)
    rc1, t1, r1, u1 = capi.fish_run(args, want_u=True)
    assert rc0 == 0 and rc1 == 0, (t0, t1)
    assert r0.its == r1.its and r1.reason > 0
    assert np.array_equal(u0, u1), _rel(u1, u0)


def test_sor3_pipelined_oracle_129():
    # the default Chebyshev/SOR multigrid at 129^3 against the oracle's
    # lexicographic SOR: iterations within one, solution within 1e-9
    args = "-fsh_dim 3 -fsh_problem manusin -da_refine 6 -pc_type mg -ksp_rtol 1e-10"
    rc, text, res, u = capi.fish_run(args, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.fish_run(args, want_u=True)
    assert orc == 0, otext
    print("sor3 129 its", res.its, "oracle", ores.its, flush=True)
    assert abs(res.its - ores.its) <= 1
    assert _rel(u, ou) < 1e-9
    assert text.splitlines()[-1] == otext.splitlines()[-1]


def test_sor3_stall_stops_the_solve():
    # a pipelined 3D sweep whose waits cannot succeed (a target no workgroup
    # reaches, a 64-poll limit): every kernel exits and the solve reports
    # PETSC_ERR_LIB; a fresh solve then runs normally
    args = "-fsh_dim 3 -fsh_problem manusin -da_refine 4 -pc_type mg -ksp_rtol 1e-8"
    rc, text, _, _ = capi.fish_run(args + " -p4g_sor_test_stall 1")
    assert rc == 76, text
    assert "pipelined SOR sweep stalled" in text
    rc, text, _, _ = capi.fish_run(args)
    assert rc == 0, text


def test_c3_513_fp64_and_mixed():
    # C3 at its full size (513^3, 135 M unknowns; the LDS-staged passes and
    # the fused residual + restriction and post-smoothing run on its finest
    # level): both precisions to the true residual 1e-10, the discretization
    # error of the noisy manufactured problem (its noise-free closed form is
    # 3.137469e-6), and the two solutions agree to 1e-9
    c3 = ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -fsh_noise_seed 0 -da_refine 8 "
          "-pc_mg_levels 8 %s -mg_levels_ksp_max_it 2 -ksp_rtol 1e-10" % MG)
    assert abs(_eig_err_3d(513) - 3.137469e-06) < 1e-11
    sol = {}
    for prec, extra in (("double", " -ksp_norm_type unpreconditioned"), ("mixed", " -p4g_precision mixed")):
        print("C3 513^3", prec, flush=True)
        with capi.Session(c3 + extra) as s:
            rep = s.solve()
            print("  its", rep.its, "relres", rep.true_relres, flush=True)
            assert rep.reason > 0
            assert rep.true_relres <= 1e-10 * 1.0001
            einf, _ = s.errors()
            assert abs(einf - 3.1329e-06) < 2e-3 * 3.1329e-06
            sol[prec] = s.solution()
    assert _rel(sol["mixed"], sol["double"]) < 1e-9
