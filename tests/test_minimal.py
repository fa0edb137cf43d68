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
