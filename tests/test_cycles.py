"""The outer Richardson KSP and the PCMG cycle variants of the reference's
cycle studies (c/ch6/study/cyclereductions.sh:19-40: -ksp_type richardson
-ksp_max_it 1 -ksp_norm_type unpreconditioned -ksp_monitor with
-pc_mg_type multiplicative|full|kaskade, -pc_mg_cycle_type v|w,
-pc_mg_multiplicative_cycles 2).

The oracle is pinned by composition with its own one-application
operators, so a wrong order of norm/update/residual, a dropped scale, a
cycle started from zero or an added (instead of overwritten) interpolation
fails:
  * the monitored Richardson norms equal |r_k| (or |B r_k|) of
    r_{k+1} = b - A x_{k+1}, x_{k+1} = x_k + s B r_k, built from PCApply and
    MatMult;
  * two multiplicative cycles equal z1 + B (r - A z1) (a cycle from an
    iterate is x + B (b - A x) for linear smoothers);
  * a two-level Galerkin K-cycle maps b = A P v to P v exactly, and is
    linear;
  * PCApplyRichardson_MG (no monitor): one or two iterations equal one or
    two cycles from zero;
  * the three-level V- and W-cycles equal the two-grid recursion
    I - B A = S (I - P C P^T A) S built from the one-level cycle (the
    smoother, S = I - B_1 A) and the two-level cycle B_c of the coarser grid,
    with C = B_c (V) and C = 2 B_c - B_c A_c B_c (W: the coarse cycle twice,
    the second from the first's iterate);
  * the two-level Galerkin cycle is S (I - P (P^T A P)^-1 P^T A) S, and the
    re-discretized one is not;
  * the three-level F-cycle is the coarsest solve of the twice-restricted b,
    then on each finer level the interpolated iterate followed by one cycle
    of that level's hierarchy from it; the Kaskade cycle the same climb with
    one level smoother application in place of each cycle.
GPU parity of the same paths is in tests/test_gpu_parity.py.
"""
import re

import numpy as np
import pytest

from oracle import oracle_py as O

MG = "-pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"
BASE2 = "-fsh_dim 2 -fsh_problem manupoly -da_refine 4 %s" % MG
BASE3 = "-fsh_dim 3 -fsh_problem manusin -da_refine 3 %s" % MG


def _monitor(text):
    return [float(l.split()[-1]) for l in text.splitlines() if "KSP Residual norm" in l]


@pytest.mark.parametrize("norm,scale", [("unpreconditioned", 1.0), ("preconditioned", 1.0),
                                        ("unpreconditioned", 0.6)])
def test_richardson_monitor_is_the_iteration(norm, scale):
    args = BASE2 + (" -ksp_type richardson -ksp_monitor -ksp_converged_reason -ksp_rtol 0 -ksp_max_it 4 "
                    "-ksp_norm_type %s -ksp_richardson_scale %g" % (norm, scale))
    rc, text, res, _ = O.fish_run(args)
    assert rc == 0, text
    got = _monitor(text)
    b = O.rhs(BASE2)
    x = np.zeros_like(b)
    r = b.copy()
    want = []
    for k in range(5):
        z = O.pc_apply(BASE2, r)
        want.append(np.linalg.norm(r if norm == "unpreconditioned" else z))
        if k == 4:
            break
        x += scale * z
        r = b - O.mat_mult(BASE2, x)
    assert len(got) == 5
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-11 * w
    assert "did not converge due to DIVERGED_ITS iterations 4" in text


def test_two_multiplicative_cycles_compose():
    rng = np.random.default_rng(3)
    r = rng.standard_normal(O.rhs(BASE3).size)
    z1 = O.pc_apply(BASE3, r)
    z2 = O.pc_apply(BASE3 + " -pc_mg_multiplicative_cycles 2", r)
    want = z1 + O.pc_apply(BASE3, r - O.mat_mult(BASE3, z1))
    assert np.max(np.abs(z2 - want)) <= 1e-12 * np.max(np.abs(want))


def test_kaskade_two_level_galerkin_is_exact_on_the_coarse_space():
    args = ("-fsh_dim 2 -fsh_problem manusin -da_refine 2 -pc_type mg -pc_mg_levels 2 -pc_mg_galerkin "
            "-pc_mg_type kaskade -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 3")
    rng = np.random.default_rng(2)
    v = rng.standard_normal(25).reshape(5, 5)
    xf = np.zeros(81)
    O.interp_add(2, 5, 5, 1, v.ravel(), xf)           # P v on the 9x9 grid
    b = O.mat_mult(args, xf)
    z = O.pc_apply(args, b)
    # coarse solve of P^T A P v gives v, the interpolation P v, and the
    # smoother leaves an exact iterate unchanged
    assert np.max(np.abs(z - xf)) <= 1e-12 * np.max(np.abs(xf))


def test_kaskade_is_linear_and_not_the_vcycle():
    args = BASE3 + " -pc_mg_type kaskade -mg_levels_ksp_max_it 4"
    rng = np.random.default_rng(5)
    r, s = rng.standard_normal((2, O.rhs(BASE3).size))
    zr, zs, zrs = O.pc_apply(args, r), O.pc_apply(args, s), O.pc_apply(args, 2 * r - 3 * s)
    assert np.max(np.abs(zrs - (2 * zr - 3 * zs))) <= 1e-12 * np.max(np.abs(zrs))
    assert np.max(np.abs(zr - O.pc_apply(BASE3 + " -mg_levels_ksp_max_it 4", r))) > 1e-3 * np.max(np.abs(zr))


@pytest.mark.parametrize("its", [1, 2])
def test_richardson_mg_shortcut_iterations_are_cycles(its):
    # no monitor: PCApplyRichardson_MG runs `its` cycles from zero (none
    # converges at rtol 1e-14 after one or two cycles)
    args = BASE3 + " -ksp_type richardson -ksp_rtol 1e-14 -ksp_max_it %d -ksp_converged_reason" % its
    rc, text, res, u = O.fish_run(args, want_u=True)
    assert rc == 0, text
    assert res.its == its and "CONVERGED_ITS iterations %d" % its in text
    b = O.rhs(BASE3)
    x = O.pc_apply(BASE3, b)
    if its == 2:
        x = x + O.pc_apply(BASE3, b - O.mat_mult(BASE3, x))
    y = -u                                  # u0 = 0 for manusin, u = u0 - y
    assert np.max(np.abs(y - x)) <= 1e-12 * np.max(np.abs(x))


def test_richardson_mg_shortcut_stops_after_presmoothing():
    # the shortcut tests |b - A x| after each finest pre-smoothing, so it
    # stops no later than the monitored loop, which tests before each cycle
    base = BASE3 + " -ksp_type richardson -ksp_norm_type unpreconditioned -ksp_rtol 1e-8 -ksp_converged_reason"
    rc1, t1, r1, _ = O.fish_run(base)
    rc2, t2, r2, _ = O.fish_run(base + " -ksp_monitor")
    assert rc1 == rc2 == 0
    assert r1.reason == r2.reason == 2
    assert r1.its <= r2.its
    assert r1.true_relres <= 1e-8


def test_three_level_v_and_w_cycles_are_the_two_grid_recursion():
    # all operators formed column by column on a 17 x 17 grid (levels 17, 9,
    # 5); the smoother's Chebyshev bounds are estimated on each level's own
    # operator, so the level smoother of every hierarchy is the same S
    def mat(f, n):
        return np.column_stack([f(e) for e in np.eye(n)])
    sm = " -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"
    fine = "-fsh_dim 2 -fsh_problem manusin -da_refine 3 -pc_type mg"
    coarse = "-fsh_dim 2 -fsh_problem manusin -da_refine 2 -pc_type mg"
    nf, nc = 17 * 17, 9 * 9
    A = mat(lambda e: O.mat_mult(fine, e), nf)
    Ac = mat(lambda e: O.mat_mult(coarse, e), nc)
    B1 = mat(lambda e: O.pc_apply(fine + " -pc_mg_levels 1" + sm, e), nf)
    Bc = mat(lambda e: O.pc_apply(coarse + " -pc_mg_levels 2" + sm, e), nc)
    P = mat(lambda e: O.interp_add(2, 9, 9, 1, e, np.zeros(nf)), nc)
    I = np.eye(nf)
    S = I - B1 @ A
    Bs = {}
    for cyc, C in (("v", Bc), ("w", 2 * Bc - Bc @ Ac @ Bc)):
        Bs[cyc] = mat(lambda e: O.pc_apply(fine + " -pc_mg_levels 3 -pc_mg_cycle_type " + cyc + sm, e), nf)
        E = I - Bs[cyc] @ A
        assert np.max(np.abs(E - S @ (I - P @ C @ P.T @ A) @ S)) <= 1e-13 * np.max(np.abs(E)), cyc
    assert np.max(np.abs(Bs["w"] - Bs["v"])) > 1e-3 * np.max(np.abs(Bs["v"]))


@pytest.mark.parametrize("fine,shape", [
    ("-fsh_dim 2 -fsh_problem manupoly -fsh_cx 0.3 -da_refine 3 -pc_type mg", (2, 9, 9, 1, 17 * 17, 9 * 9)),
    ("-fsh_dim 3 -fsh_problem manusin -da_refine 2 -pc_type mg", (3, 5, 5, 5, 9 ** 3, 5 ** 3)),
], ids=["2d_anisotropic", "3d"])
def test_two_level_galerkin_cycle_inverts_ptap(fine, shape):
    # -pc_mg_galerkin: the coarse operator is P^T A P (formed here from the
    # matrices of A and P), solved exactly on the coarse level
    dim, mxc, myc, mzc, nf, nc = shape

    def mat(f, n):
        return np.column_stack([f(e) for e in np.eye(n)])
    sm = " -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"
    A = mat(lambda e: O.mat_mult(fine, e), nf)
    B1 = mat(lambda e: O.pc_apply(fine + " -pc_mg_levels 1" + sm, e), nf)
    P = mat(lambda e: O.interp_add(dim, mxc, myc, mzc, e, np.zeros(nf)), nc)
    I = np.eye(nf)
    S = I - B1 @ A
    E2 = S @ (I - P @ np.linalg.solve(P.T @ A @ P, P.T @ A)) @ S
    for opt, exact in ((" -pc_mg_galerkin", True), ("", False)):
        E = I - mat(lambda e: O.pc_apply(fine + " -pc_mg_levels 2" + opt + sm, e), nf) @ A
        err = np.max(np.abs(E - E2)) / np.max(np.abs(E))
        assert (err <= 1e-13) if exact else (err > 1e-3), (opt, err)


def test_three_level_f_and_kaskade_cycles_are_nested_iterations():
    # PCMG full (F) cycle on levels 5^2, 9^2, 17^2: b restricted twice and
    # solved on the coarsest level, x_1 = P_1 x_0 + B_c (P_2^T b - A_1 P_1 x_0)
    # with B_c the 9^2 grid's two-level cycle, then x = P_2 x_1 + B_V (b - A
    # P_2 x_1) with B_V the three-level V-cycle (each cycle from the
    # interpolated iterate is linear: x + B (b - A x))
    def mat(f, n):
        return np.column_stack([f(e) for e in np.eye(n)])
    sm = " -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi"
    grid = lambda r: "-fsh_dim 2 -fsh_problem manusin -da_refine %d -pc_type mg" % r
    n2, n1, n0 = 17 * 17, 9 * 9, 5 * 5
    A = mat(lambda e: O.mat_mult(grid(3), e), n2)
    A1 = mat(lambda e: O.mat_mult(grid(2), e), n1)
    A0 = mat(lambda e: O.mat_mult(grid(1), e), n0)
    P2 = mat(lambda e: O.interp_add(2, 9, 9, 1, e, np.zeros(n2)), n1)
    P1 = mat(lambda e: O.interp_add(2, 5, 5, 1, e, np.zeros(n1)), n0)
    BV = mat(lambda e: O.pc_apply(grid(3) + " -pc_mg_levels 3" + sm, e), n2)
    Bc = mat(lambda e: O.pc_apply(grid(2) + " -pc_mg_levels 2" + sm, e), n1)
    BF = mat(lambda e: O.pc_apply(grid(3) + " -pc_mg_levels 3 -pc_mg_type full" + sm, e), n2)
    X0 = np.linalg.solve(A0, P1.T @ P2.T)
    X1 = P1 @ X0 + Bc @ (P2.T - A1 @ P1 @ X0)
    X2 = P2 @ X1 + BV @ (np.eye(n2) - A @ P2 @ X1)
    assert np.max(np.abs(BF - X2)) <= 1e-13 * np.max(np.abs(BF))
    assert np.max(np.abs(BF - BV)) > 1e-2 * np.max(np.abs(BV))
    # the Kaskade cycle: the same climb with one level smoother application
    # (the one-level cycle B_1 of that grid) in place of each cycle
    B1f = mat(lambda e: O.pc_apply(grid(3) + " -pc_mg_levels 1" + sm, e), n2)
    B1m = mat(lambda e: O.pc_apply(grid(2) + " -pc_mg_levels 1" + sm, e), n1)
    BK = mat(lambda e: O.pc_apply(grid(3) + " -pc_mg_levels 3 -pc_mg_type kaskade" + sm, e), n2)
    K1 = P1 @ X0 + B1m @ (P2.T - A1 @ P1 @ X0)
    K2 = P2 @ K1 + B1f @ (np.eye(n2) - A @ P2 @ K1)
    assert np.max(np.abs(BK - K2)) <= 1e-13 * np.max(np.abs(BK))
    assert np.max(np.abs(BK - BF)) > 1e-2 * np.max(np.abs(BF))
