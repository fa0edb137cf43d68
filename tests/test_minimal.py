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
minimal.test3 needs -snes_mf_operator (a matrix-free Jacobian with the
Poisson matrix as preconditioning material), which neither the oracle nor
the device implements: both refuse it.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = {t["test"]: t for t in json.load(open(os.path.join(HERE, "golden", "minimal_golden.json")))["tests"]}


@pytest.mark.parametrize("num", [1, 2, 4])
def test_oracle_golden(num):
    rc, text, res, _ = O.minimal_run(GOLDEN[num]["options"])
    assert rc == 0, text
    assert text == GOLDEN[num]["expected"]


def test_oracle_refuses_mf_operator():
    rc, text, _, _ = O.minimal_run(GOLDEN[3]["options"])
    assert rc == 56 and "snes_fd_color" in text


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
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MS_CASES))
def test_minimal_device_parity(name):
    args = MS_CASES[name] + " -snes_converged_reason -ksp_converged_reason -snes_monitor"
    print("minimal case", name, args, flush=True)
    rc, text, res, u = capi.minimal_run(args, want_u=True)
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
@pytest.mark.parametrize("num", [1, 2, 4])
def test_minimal_device_golden(num):
    args = GOLDEN[num]["options"]
    print("minimal golden", num, args, flush=True)
    rc, text, res, _ = capi.minimal_run(args)
    assert rc == 0, text
    assert _same_output(text, GOLDEN[num]["expected"]), text
    if num in (2, 4):   # no printed norms: byte for byte
        assert text == GOLDEN[num]["expected"]


@pytest.mark.gpu
def test_minimal_device_refuses_mf_operator():
    rc, text, _, _ = capi.minimal_run(GOLDEN[3]["options"])
    assert rc == 56 and "snes_mf_operator" in text
