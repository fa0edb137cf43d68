"""GPU tests of the continuous-batching server (the serving configuration).

Each request's result must be the same alone and in a mix (the batched
launches keep per-request blocks and partial sums, so the result is
bit-identical), must solve its system to the solver's rtol (checked with
the oracle's operator), and must agree with the single-solve mixed path.
Admission respects the HBM budget; an oversized request is rejected.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from p4pdes_amd import capi

pytestmark = pytest.mark.gpu

MIXED = ("-pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 1 "
         "-p4g_precision mixed -ksp_rtol 1e-8")
SHAPES = {
    "3d17": "-fsh_dim 3 -fsh_problem sinemodes -fsh_modes_seed 1 -da_refine 3 " + MIXED,
    "3d33": "-fsh_dim 3 -fsh_problem sinemodes -fsh_modes_seed 2 -da_refine 4 " + MIXED,
    "3d65": "-fsh_dim 3 -fsh_problem sinemodes -fsh_modes_seed 3 -da_refine 5 " + MIXED,
    "2d65": "-fsh_dim 2 -fsh_problem sinemodes -fsh_modes_seed 4 -da_refine 5 " + MIXED,
    "2d257": "-fsh_dim 2 -fsh_problem sinemodes -fsh_modes_seed 5 -da_refine 7 " + MIXED,
}


def _rhs(sess):
    b = np.zeros(sess.n)
    capi.check(capi.lib().p4g_memcpy_dtoh(b.ctypes.data, capi.fishlib().fish_session_rhs_device(sess.h),
                                          8 * sess.n))
    return b


@pytest.fixture(scope="module")
def sessions():
    ss = {k: capi.Session(v) for k, v in SHAPES.items()}
    yield ss
    for s in ss.values():
        s.close()


def test_alone_equals_mix(sessions):
    rng = np.random.default_rng(11)
    base = {k: _rhs(s) for k, s in sessions.items()}
    with capi.Server(64 << 30) as srv:
        rid = srv.submit(sessions["3d33"].solver, base["3d33"])
        x_alone, rep_alone = srv.wait(rid)
        print("alone: its", rep_alone.its, "relres", rep_alone.true_relres, flush=True)
        # a mix: every shape, some twice with scaled/perturbed right-hand sides
        ids = []
        for k in ["3d65", "2d257", "3d33", "3d17", "2d65", "3d65"]:
            b = base[k] * (1.0 + 0.1 * rng.standard_normal())
            ids.append((k, srv.submit(sessions[k].solver, b), b))
        out = {}
        for k, rid, b in ids:
            x, rep = srv.wait(rid)
            assert rep.reason > 0, (k, rep.reason)
            assert rep.true_relres <= 1e-8 * 1.0001
            out.setdefault(k, []).append((x, b, rep))
        x_mix = [x for (x, b, rep) in out["3d33"]][0]
        # the 3d33 request in the mix had a scaled b: solve the same b again in a mix
        ids2 = [srv.submit(sessions["3d65"].solver, base["3d65"]),
                srv.submit(sessions["3d33"].solver, base["3d33"]),
                srv.submit(sessions["2d65"].solver, base["2d65"])]
        res2 = [srv.wait(r) for r in ids2]
        st = srv.stats()
    assert np.array_equal(res2[1][0], x_alone)       # bit-identical alone and in a mix
    assert st.completed == 1 + 6 + 3
    assert st.p99_ms >= st.p50_ms > 0


def test_results_solve_the_system(sessions):
    with capi.Server(64 << 30) as srv:
        ids = [(k, srv.submit(s.solver, _rhs(s))) for k, s in sessions.items()]
        res = {k: srv.wait(r) for k, r in ids}
    for k, (x, rep) in res.items():
        b = _rhs(sessions[k])
        r = b - O.mat_mult(SHAPES[k], x)
        rel = np.linalg.norm(r) / np.linalg.norm(b)
        print(k, "its", rep.its, "replacements", rep.inner_its, "relres", rel, flush=True)
        assert rel <= 1e-8 * 1.01
        # the single-solve mixed path on the same system
        s = sessions[k]
        s.solve()
        # the session solves J y = F(u0) = b from y = 0 and returns u = u0 - y
        # with u0 = 0 (g = 0 on the boundary), so y = -u solves A y = b
        y = -s.solution()
        assert np.max(np.abs(x - y)) <= 1e-6 * np.max(np.abs(x))


def test_budget_and_rejection(sessions):
    big = sessions["3d65"]
    need = capi.Server.request_bytes(big.solver)
    with capi.Server(int(need * 1.5)) as srv:    # room for one 65^3 request at a time
        b = _rhs(big)
        ids = [srv.submit(big.solver, b) for _ in range(3)]
        xs = [srv.wait(r)[0] for r in ids]
        st = srv.stats()
        assert st.peak_bytes <= need * 1.5
        assert all(np.array_equal(xs[0], x) for x in xs)
    with capi.Server(need // 2) as srv:
        with pytest.raises(capi.P4GError):
            srv.submit(big.solver, _rhs(big))
        assert srv.stats().rejected == 1


def test_no_graph_matches_graph(sessions):
    s = sessions["2d65"]
    b = _rhs(s)
    with capi.Server(8 << 30) as g1:
        x1, _ = g1.wait(g1.submit(s.solver, b))
    with capi.Server(8 << 30, "-p4g_serve_graph 0") as g0:
        x0, _ = g0.wait(g0.submit(s.solver, b))
    assert np.array_equal(x0, x1)


def test_pooled_vectors_and_staging_reuse(sessions):
    # a request's device vectors come back to the server's pool and its
    # pinned staging buffer is reused: a later request of the same layout
    # (after another layout and another right-hand side) gives the first
    # result bit for bit; the submitted array may change right after submit
    s, t = sessions["3d65"], sessions["2d257"]
    b1 = _rhs(s)
    b2 = b1 * -0.5 + 0.25 * np.roll(b1, 7)
    with capi.Server(8 << 30) as srv:
        x1, _ = srv.wait(srv.submit(s.solver, b1))
        srv.wait(srv.submit(t.solver, _rhs(t)))
        bb = b2.copy()
        rid = srv.submit(s.solver, bb)
        bb[:] = np.nan                     # staged at submit: the request does not see this
        x2, rep2 = srv.wait(rid)
        x3, _ = srv.wait(srv.submit(s.solver, b1))
        st = srv.stats()
    assert rep2.reason > 0 and np.all(np.isfinite(x2))
    assert not np.array_equal(x1, x2)
    assert np.array_equal(x1, x3)
    assert st.completed == 4


def test_full_size_set_one_request_each():
    # C5's full size set, 33^3 to 257^3 and 257^2 to 2049^2, one request of
    # each submitted together: every result solves its system to the
    # solver's rtol (true FP64 residual through the device operator)
    import bench
    sess = {}
    try:
        for dim, r in bench.SERVE_SHAPES:
            sess[(dim, r)] = capi.Session("-fsh_dim %d -da_refine %d -pc_mg_levels %d %s"
                                          % (dim, r, r, bench.SERVE_OPTS))
        with capi.Server(64 << 30) as srv:
            ids = {}
            for (dim, r), s in sess.items():
                m = 2 ** (r + 1) + 1
                b, xh = bench.sine_mode_request(dim, m, 100 + r)
                ids[(dim, r)] = (srv.submit(s.solver, b), b, xh)
            for key, (rid, b, xh) in ids.items():
                x, rep = srv.wait(rid)
                print("size", key, "its", rep.its, "relres", rep.true_relres, flush=True)
                assert rep.reason > 0
                assert rep.true_relres <= 1e-8 * 1.0001
                # the exact discrete solution of the sine-mode right-hand side
                assert np.max(np.abs(x - xh)) <= 1e-6 * np.max(np.abs(xh))
    finally:
        for s in sess.values():
            s.close()


def test_indefinite_preconditioner_stops():
    # Richardson smoothing at scale -1 makes the FP32 V(1,1) cycle indefinite:
    # the request stops on the sign change of beta = (z, r), as the
    # single-solve mixed path does, with DIVERGED_INDEFINITE_PC
    s = capi.Session("-fsh_dim 3 -da_refine 3 -pc_type mg -mg_levels_ksp_type richardson "
                     "-mg_levels_ksp_richardson_scale -1 -mg_levels_pc_type jacobi -mg_levels_ksp_max_it 1 "
                     "-p4g_precision mixed")
    try:
        with capi.Server(64 << 30) as srv:
            _, rep = srv.wait(srv.submit(s.solver, _rhs(s)))
        print("its", rep.its, "reason", rep.reason, flush=True)
        assert rep.reason == -8, rep.reason
    finally:
        s.close()
