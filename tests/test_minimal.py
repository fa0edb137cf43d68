"""The minimal-surface equation of c/ch7/minimal.c: Newton's method with the
colored finite-difference Jacobian (-snes_fd_color), GMRES/ILU(0) (PETSc's
serial defaults) or CG, grid sequencing and the area monitor.

The oracle is pinned by
  * the reference's own outputs c/ch7/output/minimal.test1, 2 and 4 (the
    |F| history of Newton-GMRES-ILU on the 5x5 catenoid, the "Matrix is
    symmetric" lines and CG iteration counts 5 and 6 for q = 0, the grid-
    sequenced Newton counts 3, 5, 5), reproduced byte for byte;
  * the finite-difference Jacobian against central differences of the
    residual, and for q = 0 against the 5-point Laplacian it must equal;
  * GMRES against the minimizer of the preconditioned residual over the
    Krylov space, computed by numpy.
  * minimal.test3: -snes_mf_operator (the matrix-free finite-difference
    Jacobian as the operator, the registered Poisson Jacobian as
    preconditioning material) under PCMG on two processes, whose y-split
    local SOR sweeps the oracle simulates (-oracle_nranks 2; one process
    prints a different area digit);
  * the matrix-free product against the colored Jacobian's product, and for
    q = 0 (a linear residual) against the 5-point Laplacian.
"""
import re
import json
import os

import numpy as np
import pytest

from oracle import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = {t["test"]: t for t in json.load(open(os.path.join(HERE, "golden", "minimal_golden.json")))["tests"]}


def _nranks(num, dev=False):
    n = GOLDEN[num]["nprocs"]
    if n <= 1:
        return ""
    return (" -p4g_sor_blocks %d" if dev else " -oracle_nranks %d") % n


@pytest.mark.parametrize("num", [1, 2, 3, 4])
def test_oracle_golden(num):
    rc, text, res, _ = O.minimal_run(GOLDEN[num]["options"] + _nranks(num))
    assert rc == 0, text
    assert text == GOLDEN[num]["expected"]


def test_oracle_test3_needs_the_process_split():
    # one process: the same Newton counts, but an area line differs in its
    # last digit (the two processes' local SOR sweeps are a different smoother)
    rc, text, _, _ = O.minimal_run(GOLDEN[3]["options"])
    assert rc == 0, text
    assert text != GOLDEN[3]["expected"]
    got, want = text.splitlines(), GOLDEN[3]["expected"].splitlines()
    assert len(got) == len(want)
    assert [g for g in got if "Nonlinear" in g] == [w for w in want if "Nonlinear" in w]


def test_oracle_refuses_pure_mf():
    rc, text, _, _ = O.minimal_run("-snes_mf -da_refine 1")
    assert rc == 56 and "snes_mf_operator" in text


def test_mffd_product_is_the_jacobian_product():
    mx = my = 17
    c = 1.1
    u = _catenoid(mx, my, c)
    jj, ii = np.meshgrid(np.arange(my), np.arange(mx), indexing="ij")
    inner = ((ii > 0) & (ii < mx - 1) & (jj > 0) & (jj < my - 1)).ravel()
    u = u + 0.05 * np.sin(3.0 * ii.ravel() + 2.0 * jj.ravel()) * inner
    a = np.random.default_rng(3).standard_normal(mx * my)
    y = O.minimal_mffd(mx, my, 1, -0.5, 1.0, c, u, a)
    J9 = O.minimal_fdjac(mx, my, 1, -0.5, 1.0, c, u)
    Ja = np.zeros(mx * my)
    offs = _offsets(mx)
    for q in range(mx * my):
        for k, o in enumerate(offs):
            if J9[q, k] != 0.0:
                Ja[q] += J9[q, k] * a[q + o]
    assert np.max(np.abs(y - Ja)) <= 1e-5 * np.max(np.abs(Ja))
    # the zero direction gives zero (MatMult_MFFD's zeroa)
    assert not np.any(O.minimal_mffd(mx, my, 1, -0.5, 1.0, c, u, np.zeros(mx * my)))


def test_mffd_q0_is_the_laplacian():
    mx = my = 9
    u = np.random.default_rng(4).standard_normal(mx * my)
    a = np.random.default_rng(5).standard_normal(mx * my)
    y = O.minimal_mffd(mx, my, 0, 0.0, 1.0, 1.1, u, a)
    jj, ii = np.meshgrid(np.arange(my), np.arange(mx), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    want = np.zeros(mx * my)
    for q in range(mx * my):
        i, j = ii[q], jj[q]
        if i in (0, mx - 1) or j in (0, my - 1):
            want[q] = a[q]                    # F = u - g on the boundary
            continue
        s = 4.0 * a[q]                        # hx = hy: -(u_e - 2u_c + u_w) - (u_n - 2u_c + u_s)
        for o, inside in ((1, i + 1 < mx - 1), (-1, i - 1 > 0), (mx, j + 1 < my - 1), (-mx, j - 1 > 0)):
            if inside:                        # a boundary neighbour is g: no dependence on u
                s -= a[q + o]
        want[q] = s
    assert np.max(np.abs(y - want)) <= 1e-6 * np.max(np.abs(want))


def _catenoid(mx, my, c):
    x = np.linspace(0.0, 1.0, mx)[None, :]
    y = np.linspace(0.0, 1.0, my)[:, None]
    ch = np.cosh(x / c)
    return (c * ch * np.sin(np.arccos((y / c) / ch))).ravel()


def _offsets(mx):
    return [oy * mx + ox for oy in (-1, 0, 1) for ox in (-1, 0, 1)]


def test_fd_jacobian_is_the_derivative():
    mx = my = 9
    c = 1.1
    u = _catenoid(mx, my, c)
    jj, ii = np.meshgrid(np.arange(my), np.arange(mx), indexing="ij")
    inner = ((ii > 0) & (ii < mx - 1) & (jj > 0) & (jj < my - 1)).ravel()
    u = u + 0.05 * np.sin(3.0 * ii.ravel() + 2.0 * jj.ravel()) * inner   # off the solution
    J9 = O.minimal_fdjac(mx, my, 1, -0.5, 1.0, c, u)
    h = 1e-6
    off = _offsets(mx)
    worst = 0.0
    for col in range(mx * my):
        up, um = u.copy(), u.copy()
        up[col] += h
        um[col] -= h
        d = (O.minimal_residual(mx, my, 1, -0.5, 1.0, c, up) - O.minimal_residual(mx, my, 1, -0.5, 1.0, c, um)) / (2 * h)
        for row in np.nonzero(d)[0]:
            k = off.index(col - row)
            worst = max(worst, abs(J9[row, k] - d[row]))
    assert worst <= 1e-5 * np.max(np.abs(J9))


def test_fd_jacobian_q0_is_the_laplacian():
    mx = my = 9
    u = np.random.default_rng(1).standard_normal(mx * my)
    J9 = O.minimal_fdjac(mx, my, 0, 0.0, 1.0, 1.1, u)
    want = np.zeros_like(J9)
    for j in range(my):
        for i in range(mx):
            r = j * mx + i
            if i in (0, mx - 1) or j in (0, my - 1):
                want[r, 4] = 1.0
                continue
            want[r, 4] = 4.0   # 2 (hy/hx + hx/hy), hx = hy
            if i - 1 > 0:
                want[r, 3] = -1.0
            if i + 1 < mx - 1:
                want[r, 5] = -1.0
            if j - 1 > 0:
                want[r, 1] = -1.0
            if j + 1 < my - 1:
                want[r, 7] = -1.0
    assert np.max(np.abs(J9 - want)) <= 1e-6


@pytest.mark.parametrize("pc,k", [("none", 3), ("jacobi", 5)])
def test_gmres_minimizes_the_preconditioned_residual(pc, k):
    args = "-fsh_dim 2 -fsh_problem manusin -fsh_cx 0.2 -da_refine 3 -pc_type %s" % pc
    rc, text, res, u = O.fish_run(args + " -ksp_type gmres -ksp_rtol 1e-14 -ksp_max_it %d" % k, want_u=True)
    assert rc == 0, text
    y = -u                                  # u0 = 0 for manusin, u = u0 - y
    b = O.rhs(args)
    BA = lambda v: O.pc_apply(args, O.mat_mult(args, v))
    v0 = O.pc_apply(args, b)
    Q = [v0 / np.linalg.norm(v0)]
    for _ in range(k - 1):                  # Arnoldi (modified Gram-Schmidt, twice)
        w = BA(Q[-1])
        for _ in range(2):
            for q in Q:
                w = w - (q @ w) * q
        Q.append(w / np.linalg.norm(w))
    Q = np.array(Q).T
    M = np.column_stack([BA(Q[:, j]) for j in range(k)])
    cfit = np.linalg.lstsq(M, v0, rcond=None)[0]
    x = Q @ cfit
    assert np.max(np.abs(y - x)) <= 1e-9 * np.max(np.abs(x))


# ------------------------------------------------------------------ GPU

from p4pdes_amd import capi  # noqa: E402

MS_CASES = {
    "catenoid_17_ilu": "-da_refine 3 -snes_fd_color",
    "tent_33_gmres_mg": "-da_refine 4 -ms_problem tent -snes_fd_color -pc_type mg",
    "catenoid_seq3_mg_jacobi": ("-snes_grid_sequence 3 -snes_fd_color -pc_type mg -mg_levels_pc_type jacobi "
                                "-ms_catenoid_c 1.5"),
    "tent_q0_cg_ilu": "-ms_q 0 -ksp_type cg -da_refine 3 -ms_problem tent -snes_fd_color",
    "catenoid_33_none_fails": "-da_refine 4 -pc_type none -ksp_max_it 200 -snes_fd_color",
    "tent_H3_w_sor": ("-ms_problem tent -ms_tent_H 3 -da_refine 4 -pc_type mg -pc_mg_cycle_type w "
                      "-mg_levels_ksp_type richardson -mg_levels_pc_type sor -snes_fd_color"),
    "tent_seq4_mg": "-ms_problem tent -snes_grid_sequence 4 -snes_fd_color -pc_type mg",
    "catenoid_c101_seq2_restart5": ("-ms_catenoid_c 1.01 -snes_grid_sequence 2 -da_refine 1 -ksp_gmres_restart 5 "
                                    "-snes_fd_color"),
    # several 64-row strips of the pipelined SOR sweeps (65^2: a full strip
    # and a one-row strip; 129^2: two full strips and a one-row strip)
    "tent_65_mg_sor_2strips": "-da_refine 5 -ms_problem tent -snes_fd_color -pc_type mg",
    "tent_129_mg_sor_3strips": "-da_refine 6 -ms_problem tent -snes_fd_color -pc_type mg",
    "catenoid_seq6_mg_sor_3strips": "-snes_grid_sequence 6 -snes_fd_color -pc_type mg",
    "tent_H10_basic_monitor": ("-da_refine 2 -ms_problem tent -ms_tent_H 10 -snes_linesearch_type basic "
                               "-snes_fd_color -ms_monitor -ms_quaddegree 2"),
    # -snes_mf_operator: the matrix-free operator with the Poisson Jacobian as
    # preconditioning material, on one or two (simulated) processes
    "mf_catenoid_seq4_mg_2ranks": "-snes_mf_operator -pc_type mg -snes_grid_sequence 4 -oracle_nranks 2",
    "mf_tent_seq3_mg": "-snes_mf_operator -pc_type mg -snes_grid_sequence 3 -ms_problem tent",
    "mf_catenoid_17_w_richardson": ("-snes_mf_operator -pc_type mg -da_refine 3 -pc_mg_cycle_type w "
                                    "-mg_levels_ksp_type richardson"),
    "mf_catenoid_seq2_ilu_restart5": "-snes_mf_operator -snes_grid_sequence 2 -da_refine 1 -ksp_gmres_restart 5",
    # neither option: Newton on the registered Poisson Jacobian
    "poisson_jacobian_tent_H01_mg": "-ms_problem tent -ms_tent_H 0.1 -da_refine 3 -pc_type mg",
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MS_CASES))
def test_minimal_device_parity(name):
    args = MS_CASES[name] + " -snes_converged_reason -ksp_converged_reason -snes_monitor"
    # the device reads -p4g_sor_blocks, the oracle -oracle_nranks
    dev = re.sub(r"-oracle_nranks (\d+)", r"-p4g_sor_blocks \1", args)
    print("minimal case", name, dev, flush=True)
    rc, text, res, u = capi.minimal_run(dev, want_u=True)
    assert rc == 0, text
    orc, otext, ores, ou = O.minimal_run(args, want_u=True)
    assert orc == 0, otext
    assert res.reason == ores.reason and res.its == ores.its, (res.as_dict(), ores.as_dict())
    assert list(res.stage_its)[:res.stages] == list(ores.stage_its)[:ores.stages]
    assert abs(res.ksp_its_total - ores.ksp_its_total) <= max(2, ores.ksp_its_total // 20)
    if ores.reason > 0:
        assert np.max(np.abs(u - ou)) <= 1e-7 * max(1.0, np.max(np.abs(ou)))
        assert text.splitlines()[-1] == otext.splitlines()[-1]
        assert abs(res.area - ores.area) <= 1e-9 * ores.area


def _same_output(text, expected):
    """Line by line equal, except that a printed |F| may differ by one unit
    in its last (sixth) digit: the device's pow rounds differently from the
    host's in the last ulp, and the finite-difference Jacobian divides that
    by its step (h ~ 1e-8), as the host's own FMA contraction would."""
    got, want = text.splitlines(), expected.splitlines()
    if len(got) != len(want):
        return False
    for g, w in zip(got, want):
        if g == w:
            continue
        if "SNES Function norm" not in g or g.split()[:4] != w.split()[:4]:
            return False
        a, b = float(g.split()[-1]), float(w.split()[-1])
        if abs(a - b) > 2e-6 * abs(b):
            return False
    return True


@pytest.mark.gpu
@pytest.mark.parametrize("num", [1, 2, 3, 4])
def test_minimal_device_golden(num):
    args = GOLDEN[num]["options"] + _nranks(num, dev=True)
    print("minimal golden", num, args, flush=True)
    rc, text, res, _ = capi.minimal_run(args)
    assert rc == 0, text
    assert _same_output(text, GOLDEN[num]["expected"]), text
    if num in (2, 3, 4):   # no printed norms: byte for byte
        assert text == GOLDEN[num]["expected"]


@pytest.mark.gpu
def test_minimal_device_refuses_pure_mf():
    rc, text, _, _ = capi.minimal_run("-snes_mf -da_refine 1")
    assert rc == 56 and "snes_mf_operator" in text
