"""The obstacle problem (c/ch12/obstacle.c): the oracle's restatement of
PETSc's reduced-space VI Newton method, pinned on the CPU, and the device
path (p4g_vi_solve through the obstacle driver) against the oracle.

Oracle pins:
  * c/ch12/output/obstacle.test1 byte for byte (|F| history, 11 CG/ILU
    iterations per Newton step, the error line);
  * obstacle.test3 byte for byte (grid sequencing, Chebyshev/SOR multigrid
    on the inactive sets; its 3x3 grid is a one-level multigrid, the level
    smoother);
  * obstacle.test4 byte for byte (-snes_type vinewtonssls: semismooth
    Newton on the Fischer-Burmeister reformulation, grid-sequenced Newton
    counts 4, 6, 5 and the last KSP count 6); the line search's trial points
    are projected onto the bounds, and without that projection the oracle
    prints other counts (checked);
  * obstacle.test2 byte for byte (GMRES with PCASM on 4 processes: one
    subdomain per process of the 2x2 DMDA split, overlap 1, exact subdomain
    solves, the restricted form), with the 4 processes simulated
    (-oracle_nranks 4); one or two processes print other histories
    (checked);
  * an independent projected SOR solve of the discrete complementarity
    problem (numpy) on 17x17 and 33x33;
  * multigrid on the inactive sets converges in a grid-independent number
    of iterations.
The device solves each reduced Newton system in full-size vectors with the
active vertices as null rows; test_null_row_galerkin_identity checks on
the CPU that the Galerkin hierarchy of that operator equals PETSc's reduced
one (coarse inactive sets from the restricted indicator, interpolation
restricted to inactive rows and columns).
"""
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = {t["test"]: t for t in json.load(open(os.path.join(HERE, "golden", "obstacle_golden.json")))["tests"]}

AFREE = 0.697965148223374


def psi(x, y):
    r = np.sqrt(x * x + y * y)
    r0 = 0.9
    psi0 = np.sqrt(1.0 - r0 * r0)
    dpsi0 = -r0 / psi0
    return np.where(r <= r0, np.sqrt(np.maximum(1.0 - r * r, 0.0)), psi0 + dpsi0 * (r - r0))


def uexact(x, y):
    A, B = 0.680259411891719, 0.471519893402112
    r = np.sqrt(x * x + y * y)
    with np.errstate(divide="ignore"):
        return np.where(r <= AFREE, psi(x, y), -A * np.log(np.maximum(r, 1e-300)) + B)


# --------------------------------------------------------------- oracle pins

def test_oracle_golden_test1():
    t = GOLDEN[1]
    rc, text, res, _ = O.obstacle_run(t["options"])
    assert rc == 0
    assert text == t["expected"]


def test_oracle_golden_test3():
    t = GOLDEN[3]
    rc, text, res, _ = O.obstacle_run(t["options"])
    assert rc == 0
    assert text == t["expected"]
    assert res.stage_its[:4] == [1, 2, 2, 3]


def test_oracle_golden_test2_asm():
    t = GOLDEN[2]
    rc, text, res, _ = O.obstacle_run(t["options"] + " -oracle_nranks %d" % t["nprocs"])
    assert rc == 0, text
    assert text == t["expected"]


@pytest.mark.parametrize("nranks", [1, 2])
def test_oracle_test2_needs_its_four_subdomains(nranks):
    t = GOLDEN[2]
    rc, text, res, _ = O.obstacle_run(t["options"] + " -oracle_nranks %d" % nranks)
    assert rc == 0, text
    assert text != t["expected"]
    assert text.splitlines()[-1] == t["expected"].splitlines()[-1]   # the same VI solution


def test_oracle_golden_test4_ssls():
    t = GOLDEN[4]
    rc, text, res, _ = O.obstacle_run(t["options"])
    assert rc == 0, text
    assert text == t["expected"]
    assert res.stage_its[:3] == [4, 6, 5]


def test_oracle_ssls_projection_is_what_test4_pins():
    # the same method with unprojected trial points: other Newton and KSP counts
    t = GOLDEN[4]
    rc, text, res, _ = O.obstacle_run(t["options"] + " -oracle_ls_project 0")
    assert rc == 0, text
    assert text != t["expected"]
    assert text.splitlines()[-1] == t["expected"].splitlines()[-1]   # the same VI solution


@pytest.mark.parametrize("num", [2, 4])
def test_oracle_reproduces_error_lines_of_other_solvers(num):
    # test2: GMRES + ASM on 4 ranks, test4: the semismooth method with grid
    # sequencing -- both end on the 9x9 VI solution, which is unique
    want = GOLDEN[num]["expected"].splitlines()
    rc, text, res, _ = O.obstacle_run("-da_refine 2 -ksp_rtol 1e-12 -snes_rtol 1e-12")
    assert rc == 0
    assert text.splitlines()[-1] == want[-1]
    assert res.mx == res.my == 9


def _operator(m):
    """A and F(0) of Poisson2DJacobianLocal/FunctionLocal on (-2,2)^2, g = u_exact, f = 0."""
    h = 4.0 / (m - 1)
    x = -2.0 + h * np.arange(m)
    X, Y = np.meshgrid(x, x)          # [j, i]
    n = m * m
    bnd = np.zeros((m, m), bool)
    bnd[0, :] = bnd[-1, :] = bnd[:, 0] = bnd[:, -1] = True
    g = uexact(X, Y)
    A = np.zeros((n, n))
    F0 = np.zeros(n)
    idx = np.arange(n).reshape(m, m)
    for j in range(m):
        for i in range(m):
            q = idx[j, i]
            A[q, q] = 4.0
            if bnd[j, i]:
                F0[q] = -4.0 * g[j, i]
                continue
            for dj, di in ((0, -1), (0, 1), (-1, 0), (1, 0)):
                if bnd[j + dj, i + di]:
                    F0[q] -= g[j + dj, i + di]
                else:
                    A[q, idx[j + dj, i + di]] = -1.0
    return A, F0, psi(X, Y).ravel(), X, Y


@pytest.mark.parametrize("refine", [3, 4])
def test_oracle_vi_solution_matches_projected_sor(refine):
    m = 2 ** (refine + 1) + 1
    A, F0, xl, X, Y = _operator(m)
    colors = ((np.add.outer(np.arange(m), np.arange(m)) % 2).ravel() == 0)
    omega = 2.0 / (1.0 + np.sin(np.pi / (m - 1)))
    u = np.maximum(np.zeros(m * m), xl)
    for sweep in range(20000):
        du = 0.0
        for c in (colors, ~colors):
            r = A[c] @ u + F0[c]
            new = np.maximum(xl[c], u[c] - omega * r / 4.0)
            du = max(du, np.max(np.abs(new - u[c])))
            u[c] = new
        if du < 1e-15:
            break
    rc, text, res, uo = O.obstacle_run("-da_refine %d -ksp_rtol 1e-12 -snes_rtol 1e-12" % refine, want_u=True)
    assert rc == 0
    assert np.max(np.abs(uo - u)) < 1e-10
    F = A @ u + F0
    assert res.active == int(np.sum((u <= xl + 1e-8) & (F > 0)))


def test_oracle_vi_multigrid_iterations_grid_independent():
    its = []
    for refine in (4, 5, 6):
        rc, text, res, _ = O.obstacle_run("-da_refine %d -pc_type mg -mg_levels_pc_type jacobi "
                                          "-ksp_converged_reason" % refine)
        assert rc == 0
        k = [int(v) for v in re.findall(r"Linear solve converged due to \w+ iterations (\d+)", text)]
        assert k and max(k) <= 8
        its.append(max(k))
    assert max(its) - min(its) <= 2


def _q1(mc):
    mf = 2 * (mc - 1) + 1
    P = np.zeros((mf * mf, mc * mc))
    w = lambda f: [(f // 2, 1.0)] if f % 2 == 0 else [(f // 2, 0.5), (f // 2 + 1, 0.5)]
    for j in range(mf):
        for i in range(mf):
            for J, a in w(j):
                for I, b in w(i):
                    P[j * mf + i, J * mc + I] += a * b
    return P


def test_null_row_galerkin_identity():
    rng = np.random.default_rng(5)
    m, mc = 17, 9
    A, _, _, X, Y = _operator(m)
    act = (np.sqrt(X * X + Y * Y) < 0.8).ravel() | (rng.random(m * m) < 0.05)
    act[0] = True   # a boundary vertex may be active too
    I = ~act
    An = A * np.outer(I, I)                 # null rows and columns at active vertices
    P = _q1(mc)
    Ac = P.T @ An @ P
    # PETSc's reduced hierarchy: coarse inactive = restricted indicator != 0
    Ic = (P.T @ I.astype(float)) != 0.0
    Pr = P[np.ix_(I, Ic)]
    Ared = Pr.T @ A[np.ix_(I, I)] @ Pr
    assert np.allclose(Ac[np.ix_(Ic, Ic)], Ared, rtol=0, atol=1e-13)
    assert np.all(Ac[~Ic] == 0.0) and np.all(Ac[:, ~Ic] == 0.0)
    # an inactive fine vertex interpolates only from inactive coarse vertices
    assert np.all(P[np.ix_(I, ~Ic)] == 0.0)


# ---------------------------------------------------------- device parity

CASES = {
    "65_gmres_ilu": "-da_refine 5 -ksp_type gmres",
    "seq3_gmres_mg": "-snes_grid_sequence 3 -ksp_type gmres -pc_type mg -mg_levels_pc_type jacobi",
    "9_mg": "-da_refine 2 -pc_type mg -mg_levels_pc_type jacobi",
    "33_mg": "-da_refine 4 -pc_type mg -mg_levels_pc_type jacobi",
    "129_mg": "-da_refine 6 -pc_type mg -mg_levels_pc_type jacobi",
    "257_mg_richardson": "-da_refine 7 -pc_type mg -mg_levels_ksp_type richardson -mg_levels_pc_type jacobi "
                         "-mg_levels_ksp_richardson_scale 0.6",
    "65_w_cycle": "-da_refine 5 -pc_type mg -mg_levels_pc_type jacobi -pc_mg_cycle_type w",
    "33_three_levels": "-da_refine 4 -pc_type mg -pc_mg_levels 3 -mg_levels_pc_type jacobi "
                       "-mg_levels_ksp_max_it 3",
    "65_unpreconditioned": "-da_refine 5 -pc_type mg -mg_levels_pc_type jacobi -ksp_norm_type unpreconditioned",
    "33_jacobi": "-da_refine 4 -pc_type jacobi",
    "33_none_basic": "-da_refine 4 -pc_type none -snes_linesearch_type basic",
    "seq3": "-snes_grid_sequence 3 -pc_type mg -mg_levels_pc_type jacobi",
    "seq4_from_9": "-da_refine 2 -snes_grid_sequence 4 -pc_type mg -mg_levels_pc_type jacobi",
    # without grid sequencing the free boundary moves about one cell per
    # Newton step: 513x513 from u = 0 reaches -snes_max_it; sequence instead
    "seq5_to_513": "-da_refine 3 -snes_grid_sequence 5 -pc_type mg -mg_levels_pc_type jacobi",
    "129_max_it": "-da_refine 6 -pc_type mg -mg_levels_pc_type jacobi -snes_max_it 5",
    "65_sor": "-da_refine 5 -pc_type mg",
    # pipelined strip SOR on the VI Galerkin levels: two full 64-row strips
    # and a one-row strip at 129^2; a grid sequence ending at 257^2
    "129_sor": "-da_refine 6 -pc_type mg",
    "seq5_to_257_sor": "-da_refine 2 -snes_grid_sequence 5 -pc_type mg",
    "65_ilu": "-da_refine 5 -ksp_rtol 1e-8",
    "seq3_ilu": "-snes_grid_sequence 3",
    "seq4_sor_w": "-snes_grid_sequence 4 -pc_type mg -pc_mg_cycle_type w",
    # the semismooth method (SNESVINEWTONSSLS)
    "ssls_33_ilu": "-snes_type vinewtonssls -da_refine 4",
    "ssls_seq4_gmres": "-snes_type vinewtonssls -snes_grid_sequence 4 -ksp_type gmres",
    "ssls_17_none": "-snes_type vinewtonssls -da_refine 3 -pc_type none",
    # PCASM (restricted, exact subdomain solves) over simulated processes
    "asm_33_gmres_4ranks": "-da_refine 4 -ksp_type gmres -pc_type asm -sub_pc_type lu -oracle_nranks 4",
    "asm_17_cg_2ranks_overlap2": ("-da_refine 3 -pc_type asm -sub_pc_type lu -pc_asm_overlap 2 "
                                  "-oracle_nranks 2"),
    "asm_seq3_gmres_6ranks": ("-snes_grid_sequence 3 -ksp_type gmres -pc_type asm -sub_pc_type lu "
                              "-oracle_nranks 6"),
}
SHOW = " -snes_monitor_short -ksp_converged_reason -snes_converged_reason"


def _ksp_its(text):
    return [int(v) for v in re.findall(r"Linear solve (?:did not )?converge[d]? due to \w+ iterations (\d+)", text)]


def _fnorms(text):
    out = []
    for line in text.splitlines():
        mm = re.search(r"SNES Function norm (.*)$", line)
        if mm:
            v = mm.group(1)
            out.append(0.0 if v.startswith("<") else float(v))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_obstacle_device_parity(case):
    from p4pdes_amd import capi
    args = CASES[case] + SHOW
    dev = re.sub(r"-oracle_nranks (\d+)", r"-p4g_ranks \1", args)   # the device's simulated processes
    print("obstacle case", case, dev, flush=True)
    rc, text, res, u = capi.obstacle_run(dev, want_u=True)
    assert rc == 0, text
    rco, texto, reso, uo = O.obstacle_run(args, want_u=True)
    assert rco == 0, texto
    assert res.its == reso.its and res.reason == reso.reason
    assert list(res.stage_its)[:res.stages] == list(reso.stage_its)[:reso.stages]
    ko, kd = _ksp_its(texto), _ksp_its(text)
    assert len(kd) == len(ko)
    assert all(abs(a - b) <= 1 for a, b in zip(kd, ko)), (kd, ko)
    fo, fd = _fnorms(texto), _fnorms(text)
    assert len(fd) == len(fo)
    scale = max(fo)
    for a, b in zip(fd, fo):
        if b > 1e-7 * scale:
            assert abs(a - b) <= 1e-5 * b, (fd, fo)   # printed with %g: 6 digits
    assert abs(res.fnorm0 - reso.fnorm0) <= 1e-12 * reso.fnorm0
    assert np.max(np.abs(u - uo)) <= 1e-9 * np.max(np.abs(uo))
    assert res.active == reso.active
    assert text.splitlines()[-1] == texto.splitlines()[-1]   # the error line


@pytest.mark.gpu
def test_obstacle_device_golden_test1_default_ilu():
    # obstacle.test1 as the reference runs it: CG with PETSc's default ILU(0)
    # on each reduced Newton system, byte for byte
    from p4pdes_amd import capi
    t = GOLDEN[1]
    print("obstacle golden 1 default ILU", flush=True)
    rc, text, res, _ = capi.obstacle_run(t["options"])
    assert rc == 0, text
    assert text == t["expected"]


@pytest.mark.gpu
def test_obstacle_device_golden_test3_default_smoother():
    # obstacle.test3 as the reference runs it (-pc_type mg with PETSc's
    # default Chebyshev/SOR smoother on the inactive-set Galerkin levels),
    # byte for byte
    from p4pdes_amd import capi
    t = GOLDEN[3]
    print("obstacle golden 3 default smoother", flush=True)
    rc, text, res, _ = capi.obstacle_run(t["options"])
    assert rc == 0, text
    assert text == t["expected"]


@pytest.mark.gpu
def test_obstacle_device_golden_test2_asm():
    # obstacle.test2 as the reference runs it on 4 processes (GMRES with
    # PCASM, one subdomain per process), byte for byte
    from p4pdes_amd import capi
    t = GOLDEN[2]
    print("obstacle golden 2 asm", flush=True)
    rc, text, res, _ = capi.obstacle_run(t["options"] + " -p4g_ranks %d" % t["nprocs"])
    assert rc == 0, text
    assert text == t["expected"]


@pytest.mark.gpu
def test_obstacle_device_golden_test4_ssls():
    # obstacle.test4 as the reference runs it (-snes_type vinewtonssls, CG
    # with PETSc's default ILU(0) on the semismooth Jacobian), byte for byte
    from p4pdes_amd import capi
    t = GOLDEN[4]
    print("obstacle golden 4 ssls", flush=True)
    rc, text, res, _ = capi.obstacle_run(t["options"])
    assert rc == 0, text
    assert text == t["expected"]


@pytest.mark.gpu
@pytest.mark.parametrize("num", [1, 3])
def test_obstacle_device_golden_lines(num):
    # the reference's commands with Jacobi-smoothed multigrid in place of
    # their own PCs (test1: PETSc's default ILU(0); test3: Chebyshev/SOR,
    # both run byte for byte in their own tests): the Newton history and the
    # report do not depend on the preconditioner
    from p4pdes_amd import capi
    t = GOLDEN[num]
    args = t["options"] + " -pc_type mg -mg_levels_pc_type jacobi"
    print("obstacle golden", num, args, flush=True)
    rc, text, res, _ = capi.obstacle_run(args)
    assert rc == 0, text
    strip = lambda s: [re.sub(r"last KSP iters = \d+", "", l) for l in s.splitlines()
                       if "Linear solve" not in l]
    got, want = strip(text), strip(t["expected"])
    if num == 3:
        # the coarser grids' Newton counts depend on how accurately the
        # smoother lets each linear solve reach rtol 1e-5; the finest grid's
        # count, the report and the errors do not
        got, want = got[-3:], want[-3:]
    assert got == want


@pytest.mark.gpu
def test_obstacle_sor_stall_stops_the_solve():
    # A pipelined SOR sweep whose waits cannot succeed (the host gives every
    # wait a target no strip reaches and a 64-poll limit, so every kernel
    # still exits): the solve stops after its first linear solve with
    # PETSC_ERR_LIB, and the next solve on a fresh solver runs normally.
    from p4pdes_amd import capi
    print("obstacle stall test", flush=True)
    rc, text, _, _ = capi.obstacle_run("-da_refine 5 -pc_type mg -p4g_sor_test_stall 1")
    assert rc == 76, text
    assert "pipelined SOR sweep stalled" in text
    rc, text, res, _ = capi.obstacle_run("-da_refine 5 -pc_type mg")
    assert rc == 0, text


@pytest.mark.gpu
def test_obstacle_dump_binary(tmp_path):
    # -obs_dump_binary (c/ch12/obstacle.c:80-83,137-146): the solution, then
    # the obstacle psi, as two PETSc binary Vecs in one file, announced by the
    # reference's line before the report
    from p4pdes_amd import capi
    path = tmp_path / "dump.dat"
    args = "-da_refine 3 -pc_type mg -mg_levels_pc_type jacobi -obs_dump_binary %s" % path
    print("obstacle dump", args, flush=True)
    rc, text, res, u = capi.obstacle_run(args, want_u=True)
    assert rc == 0, text
    lines = text.splitlines()
    assert lines[-3] == "writing u,psi in binary format to %s ..." % path
    assert lines[-2].startswith("done on 17 x 17 grid")
    ud = capi.vec_load_binary(path, 0)
    pd = capi.vec_load_binary(path, 1)
    assert np.array_equal(ud, u)
    xs = -2.0 + 4.0 * np.arange(17) / 16
    X, Y = np.meshgrid(xs, xs)
    assert np.allclose(pd, psi(X, Y).ravel(), rtol=0, atol=1e-15)
    assert np.all(ud >= pd - 1e-12)
