"""Benchmark: 3D Poisson DOF/s to rtol 1e-10 with CG + geometric multigrid on
one MI355X, mixed precision vs FP64, p50/p99 solve latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-513]

A step is one KSP solve (CG + V(1,1) Chebyshev/Jacobi multigrid, from x = 0)
of the 257^3 FP64 system (16,974,593 DOF = mx*my*mz, the reference's global
Vec size), with F(u0) already resident in HBM; set-up (callbacks,
hierarchy, eigen-estimates, coarse factorization) is outside the timed
region, as KSPSolve is timed apart from SNESSolve in the reference's
studies.  Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MG = ("-pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi "
      "-mg_coarse_pc_type cholesky")
C2 = ("-fsh_dim 3 -fsh_problem manusin -da_refine 7 -pc_mg_levels 7 %s "
      "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG)
# C3: 513^3, f = manufactured + N(0, 1e-3) noise, Chebyshev degree 2; both
# precisions converge on the true FP64 residual |b - A x|_2 / |b|_2 <= 1e-10
C3 = ("-fsh_dim 3 -fsh_problem manusin -fsh_noise_sigma 1e-3 -fsh_noise_seed 0 -da_refine 8 "
      "-pc_mg_levels 8 %s -mg_levels_ksp_max_it 2 -ksp_rtol 1e-10" % MG)
# C4: 257^3 vertices, cell coefficient k = exp(G), G a seeded Gaussian random field (random
# Fourier features, sigma 1, correlation length 0.1); manufactured solution sin sin sin
C4 = ("-fsh_dim 3 -fsh_problem varcoef -da_refine 7 -pc_mg_levels 7 %s "
      "-mg_levels_ksp_max_it 1 -ksp_rtol 1e-10" % MG)
HBM_PEAK_GBS = 8000.0
METRIC = "3D Poisson DOF/s to rtol 1e-10, mixed vs FP64 MG-CG; p50/p99 solve latency"


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def time_solves(sess, steps, warmup):
    from p4pdes_amd import capi
    for _ in range(warmup):
        sess.solve()
    capi.check(capi.lib().p4g_device_synchronize())
    lat = []
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        reps.append(sess.solve())
        lat.append((time.perf_counter() - t) * 1e3)
    capi.check(capi.lib().p4g_device_synchronize())
    total = time.perf_counter() - t0
    return total, lat, reps


TRAFFIC_FILE = "profiles/hbm_traffic_fp64_257cubed.json"


def measured_traffic(cls_name):
    """HBM bytes per finest-level launch of a kernel class on the 257^3
    problem, from the calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    recorded in TRAFFIC_FILE (None when absent)."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("kernels", {}).get(cls_name)


def kernel_roofline(sess, cls):
    """Average duration and algorithmic bytes per launch of kernel class
    `cls` on the finest level, from HIP events recorded around each launch
    on the solver's stream during the solves since the last reset."""
    n, ms, by = sess.kernel_stats(cls, fine_only=True)
    if n == 0 or ms <= 0:
        return None
    return {"launches": n, "avg_ms": ms / n, "bytes_per_launch": by / n,
            "gbs": (by / n) / (ms / n * 1e-3) / 1e9}


def stencil_apply_roofline(sess, reps=20):
    """y = A x (2 FP64 passes) on the session's finest grid, HIP-event timed."""
    from p4pdes_amd import capi
    L = capi.lib()
    nb = 8 * sess.n
    x = L.p4g_device_malloc(nb)
    y = L.p4g_device_malloc(nb)
    try:
        rhs = capi.fishlib().fish_session_rhs_device(sess.h)
        capi.check(L.p4g_mat_mult_device(sess.solver, rhs, x))   # warm + non-trivial input
        sess.reset_stats()
        sess.set_timing(capi.K_APPLY, True)
        for _ in range(reps):
            capi.check(L.p4g_mat_mult_device(sess.solver, x, y))
        sess.set_timing(capi.K_APPLY, False)
        r = kernel_roofline(sess, capi.K_APPLY)
    finally:
        L.p4g_device_free(x)
        L.p4g_device_free(y)
    return r


def cpu_baseline(args, nsolves=16):
    """The oracle (C restatement, CSR operator as PETSc's AIJ) on the host
    cores: nsolves KSP solves of the same system after one warm-up (about
    10 s of CPU work at 0.6 s per solve); the median solve time."""
    from oracle import oracle_py
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(16, ncores))
    times, res = oracle_py.fish_time(args, nthreads=threads, warmup=1, nsolves=nsolves)
    _, mx, my, mz = oracle_py.grid_shape(args)
    n = mx * my * mz
    med = sorted(times)[len(times) // 2]
    return {"value": n / (med * 1e-3), "unit": "DOF/s", "cores": threads, "kind": "port",
            "sample": "%d KSP solves of the full 257^3 system after one warm-up (median %.2f s, total "
                      "%.1f s), oracle CSR CG + V(1,1) Chebyshev/Jacobi multigrid, %d OpenMP threads, "
                      "%d its" % (nsolves, med * 1e-3, sum(times) * 1e-3, threads, res.its)}


SERVE_SHAPES = [(3, 4), (3, 5), (3, 6), (3, 7), (2, 7), (2, 9), (2, 10)]   # (dim, refine)
SERVE_OPTS = ("-fsh_problem zero -pc_type mg -mg_levels_ksp_type chebyshev -mg_levels_pc_type jacobi "
              "-mg_levels_ksp_max_it 1 -p4g_precision mixed -ksp_rtol 1e-8")


def sine_mode_request(dim, m, seed):
    """b = measure * f at interior vertices for f = sum of 8 sine modes with
    N(0,1) amplitudes and wavenumbers 1..8 (seeded); the exact discrete
    solution is sum a_k mode_k / lambda_h(k)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    h = 1.0 / (m - 1)
    xs = np.arange(m) * h
    f = np.zeros((m,) * dim)
    xh = np.zeros((m,) * dim)
    for _ in range(8):
        a = rng.standard_normal()
        ks = rng.integers(1, 9, size=dim)
        vecs = [np.sin(k * np.pi * xs) for k in ks]
        lam = sum(4.0 / h ** 2 * np.sin(k * np.pi * h / 2) ** 2 for k in ks)
        mode = vecs[0]
        for v in vecs[1:]:
            mode = np.multiply.outer(v, mode)    # z (slowest) ... x (fastest)
        f += a * mode
        xh += (a / lam) * mode
    b = (h ** dim) * f
    for ax in range(dim):        # boundary rows: b = 0 (x = 0 there)
        sl = [slice(None)] * dim
        sl[ax] = 0
        b[tuple(sl)] = 0.0
        sl[ax] = m - 1
        b[tuple(sl)] = 0.0
    return b.ravel(), xh.ravel()


def serving(rate=40.0, nreq=80, seed=2026, budget=64 << 30):
    """Poisson-arrival stream of mixed-size solves through the
    continuous-batching server; p50/p99 latency and throughput."""
    import numpy as np
    from p4pdes_amd import capi
    rng = np.random.default_rng(seed)
    sess = {}
    for dim, r in SERVE_SHAPES:
        sess[(dim, r)] = capi.Session("-fsh_dim %d -da_refine %d -pc_mg_levels %d %s"
                                      % (dim, r, r, SERVE_OPTS))
    shapes = [SERVE_SHAPES[i] for i in rng.integers(0, len(SERVE_SHAPES), size=nreq)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    reqs = []
    for q, (dim, r) in enumerate(shapes):
        m = 2 ** (r + 1) + 1
        b, xh = sine_mode_request(dim, m, seed * 1000 + q)
        reqs.append((dim, r, m, b, xh if q < 12 else None))
    out = []
    with capi.Server(budget) as srv:
        t0 = time.perf_counter()
        ids = []
        for q, (dim, r, m, b, _) in enumerate(reqs):
            delay = arrivals[q] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(sess[(dim, r)].solver, b))
        for q, rid in enumerate(ids):
            out.append(srv.wait(rid))
        st = srv.stats()
    errs = []
    for q, (dim, r, m, b, xh) in enumerate(reqs):
        if xh is not None:
            errs.append(float(np.max(np.abs(out[q][0] - xh)) / np.max(np.abs(xh))))
    for s in sess.values():
        s.close()
    its = [rep.its for _, rep in out]
    return {"p50_ms": st.p50_ms, "p99_ms": st.p99_ms, "mean_ms": st.mean_ms, "max_ms": st.max_ms,
            "queue_p50_ms": st.queue_p50_ms, "throughput_dof_per_s": st.dof_per_s,
            "requests": int(st.completed), "rate_per_s": rate, "wall_s": st.wall_s,
            "rounds": int(st.rounds), "graph_captures": int(st.graph_captures),
            "peak_hbm_gb": st.peak_bytes / 2 ** 30, "its_median": float(np.median(its)),
            "max_rel_err_vs_exact_discrete": max(errs) if errs else None,
            "all_converged": all(rep.reason > 0 for _, rep in out),
            "workload": "Poisson arrivals (seeded) at %.0f/s of %d solves, sizes uniform over 33^3, 65^3, "
                        "129^3, 257^3, 257^2, 1025^2, 2049^2; RHS = 8 random sine modes, N(0,1) "
                        "amplitudes; zero Dirichlet; mixed precision to rtol 1e-8; V(1,1) "
                        "Chebyshev/Jacobi to 5^d; HBM budget %d GiB" % (rate, nreq, budget >> 30)}


COPY_BYTES = 1 << 30


def profile_workload():
    """Workload for hardware-counter passes: the 257^3 solve, then a device
    copy of COPY_BYTES (read COPY_BYTES, write COPY_BYTES) whose counter
    readings calibrate the counters' bytes."""
    from p4pdes_amd import capi
    with capi.Session(C2) as s:
        for _ in range(2):
            s.solve()
    L = capi.lib()
    src = L.p4g_device_malloc(COPY_BYTES)
    dst = L.p4g_device_malloc(COPY_BYTES)
    if not src or not dst:
        raise RuntimeError("device allocation failed")
    for _ in range(3):
        capi.check(L.p4g_memcpy_dtod(dst, src, COPY_BYTES))
    capi.check(L.p4g_device_synchronize())
    L.p4g_device_free(src)
    L.p4g_device_free(dst)
    print("profile workload done", flush=True)


OBSTACLE = "-da_refine 2 -snes_grid_sequence 8 -pc_type mg -mg_levels_pc_type jacobi"


def obstacle(with_oracle):
    """The obstacle problem from 9x9 through 8 grid refinements to 2049x2049:
    reduced-space VI Newton per grid, CG + Galerkin V(2,2) Chebyshev/Jacobi on
    the inactive sets.  Host wall time of the whole sequence (set-up, callbacks
    and transfers included; the iterate crosses PCIe once per grid), and the
    oracle's time for the same run."""
    from p4pdes_amd import capi
    capi.obstacle_run(OBSTACLE.replace("-snes_grid_sequence 8", "-snes_grid_sequence 3"))   # warm-up
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc, text, res, _ = capi.obstacle_run(OBSTACLE)
        walls.append(time.perf_counter() - t0)
        if rc:
            return {"error": text}
    wall = sorted(walls)[1]
    dof = res.mx * res.my
    o = {"workload": "obstacle problem of c/ch12/obstacle.c, grid-sequenced 9^2 -> 2049^2 (%d DOF on the "
                     "final grid), SNESVINEWTONRSLS per grid, CG + Galerkin multigrid on the inactive sets "
                     "(V(2,2) Chebyshev/Jacobi)" % dof,
         "options": OBSTACLE, "wall_s": wall, "dof_per_s": dof / wall, "stage_its": res.as_dict()["stage_its"],
         "ksp_its_final_grid": res.ksp_its_total, "report": text.splitlines()[-2:]}
    if with_oracle:
        from oracle import oracle_py
        t0 = time.perf_counter()
        rco, texto, reso, _ = oracle_py.obstacle_run(OBSTACLE)
        o["oracle_wall_s"] = time.perf_counter() - t0
        o["oracle_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o["oracle_report_equal"] = texto.splitlines()[-2:] == o["report"]
    return o


OBSTACLE_CLUSTER = ("-da_grid_x 33 -da_grid_y 33 -snes_grid_sequence 6 -snes_converged_reason -ksp_converged_reason "
                    "-pc_type mg")


def obstacle_cluster(with_oracle):
    """The obstacle problem as c/ch8/cluster.sh:78-82 runs it: 33^2 base grid, six
    refinements to 2049^2, PETSc's default Chebyshev/SOR level smoothing (the
    SOR sweeps as pipelined strips).  Host wall time of the whole sequence,
    and the oracle's time for the same run."""
    from p4pdes_amd import capi
    capi.obstacle_run(OBSTACLE_CLUSTER.replace("-snes_grid_sequence 6", "-snes_grid_sequence 1"))   # warm-up
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc, text, res, _ = capi.obstacle_run(OBSTACLE_CLUSTER)
        walls.append(time.perf_counter() - t0)
        if rc:
            return {"error": text}
    wall = sorted(walls)[1]
    dof = res.mx * res.my
    o = {"workload": "obstacle problem of c/ch12/obstacle.c as c/ch8/cluster.sh runs it, grid-sequenced 33^2 -> "
                     "2049^2 (%d DOF on the final grid), CG + Galerkin multigrid on the inactive sets with "
                     "Chebyshev/SOR smoothing" % dof,
         "options": OBSTACLE_CLUSTER, "wall_s": wall, "dof_per_s": dof / wall,
         "stage_its": res.as_dict()["stage_its"], "report": text.splitlines()[-2:]}
    if with_oracle:
        from oracle import oracle_py
        t0 = time.perf_counter()
        rco, texto, reso, _ = oracle_py.obstacle_run(OBSTACLE_CLUSTER)
        o["oracle_wall_s"] = time.perf_counter() - t0
        o["oracle_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o["oracle_report_equal"] = texto.splitlines()[-2:] == o["report"]
    return o


MINIMAL = ("-da_grid_x 33 -da_grid_y 33 -snes_grid_sequence 6 -snes_fd_color -snes_converged_reason "
           "-ksp_converged_reason -pc_type mg")


def minimal(with_oracle):
    """The minimal-surface equation as c/ch8/cluster.sh:66-70 runs it: catenoid
    boundary values, 33^2 base grid, 6 grid refinements to 2049^2, Newton
    with colored finite-difference Jacobians per grid, GMRES + multigrid over
    the finite-difference Jacobians of every level.  Host wall time of the
    whole sequence (set-up, callbacks and transfers included), and the
    oracle's time for the same run through 1025^2 (5 refinements: about
    15 s of one core; the full run takes the oracle about 4 times that)."""
    from p4pdes_amd import capi
    capi.minimal_run(MINIMAL.replace("-snes_grid_sequence 6", "-snes_grid_sequence 2"))   # warm-up
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc, text, res, _ = capi.minimal_run(MINIMAL)
        walls.append(time.perf_counter() - t0)
        if rc:
            return {"error": text}
    wall = sorted(walls)[1]
    dof = res.mx * res.my
    o = {"workload": "minimal-surface equation of c/ch7/minimal.c (catenoid), grid-sequenced 33^2 -> 2049^2 "
                     "(%d DOF on the final grid), Newton per grid with colored finite-difference Jacobians, "
                     "GMRES + multigrid over the level Jacobians (Chebyshev/SOR)" % dof,
         "options": MINIMAL, "wall_s": wall, "dof_per_s": dof / wall, "stage_its": res.as_dict()["stage_its"],
         "ksp_its_total": res.ksp_its_total, "report": text.splitlines()[-1]}
    if with_oracle:
        from oracle import oracle_py
        sample = MINIMAL.replace("-snes_grid_sequence 6", "-snes_grid_sequence 5")
        rc5, text5, res5, _ = capi.minimal_run(sample)
        t0 = time.perf_counter()
        rco, texto, reso, _ = oracle_py.minimal_run(sample)
        o["oracle_sample"] = "the same options through 1025^2 (-snes_grid_sequence 5)"
        o["oracle_sample_wall_s"] = time.perf_counter() - t0
        o["oracle_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o["oracle_report_equal"] = texto.splitlines()[-1] == text5.splitlines()[-1]
        o["oracle_stage_its_equal"] = reso.as_dict()["stage_its"] == res5.as_dict()["stage_its"]
    return o


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-513", action="store_true", help="skip the 513^3 mixed/FP64 runs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-obstacle", action="store_true", help="skip the obstacle (VI) solve")
    ap.add_argument("--no-minimal", action="store_true", help="skip the minimal-surface solve")
    ap.add_argument("--profile-workload", action="store_true",
                    help="only run the counter-collection workload: two 257^3 solves and a 1 GiB "
                         "device copy (the calibration copy of known size); prints no JSON")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU z-slab decomposition is a design, not built)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    from p4pdes_amd import capi
    if capi.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if a.profile_workload:
        profile_workload()
        return

    out = {"metric": METRIC, "unit": "DOF/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (manufactured sine right-hand side, closed-form solution)"}

    # ---- headline: C2 257^3 FP64 ----------------------------------------
    print("bench: setting up 257^3", file=sys.stderr, flush=True)
    with capi.Session(C2) as s:
        n = s.n
        s.solve()
        # roofline of the dominant kernel: time every class once (untimed run
        # is unaffected), then event-time the dominant class in the timed run
        s.reset_stats()
        for c in range(capi.K_COUNT):
            s.set_timing(c, True)
        s.solve()
        by_cls = {c: s.kernel_stats(c)[1] for c in range(capi.K_COUNT)}
        solve_bytes = sum(s.kernel_stats(c)[2] for c in range(capi.K_COUNT))   # algorithmic, one solve
        for c in range(capi.K_COUNT):
            s.set_timing(c, False)
        dominant = max((c for c in by_cls if c not in (capi.K_OTHER, capi.K_REDUCE)),
                       key=lambda c: by_cls[c])
        breakdown = {capi.KERNEL_NAMES[c]: round(by_cls[c], 4) for c in by_cls}
        s.reset_stats()
        s.set_timing(dominant, True, fine_only=True)   # one event pair per cycle
        print("bench: timing %d solves" % a.steps, file=sys.stderr, flush=True)
        total, lat, reps = time_solves(s, a.steps, a.warmup)
        s.set_timing(dominant, False)
        roof = kernel_roofline(s, dominant)
        einf, e2h = s.errors()
        apply257 = stencil_apply_roofline(s)
        rep = reps[-1]
    ms = total / a.steps * 1e3
    out["value"] = n / (ms * 1e-3)
    out["ms_per_step"] = ms
    out["latency_ms"] = {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99)}
    out["config"] = {"workload": "3D Poisson 257^3 vertex grid (16,974,593 DOF), 7-point stencil, "
                                 "u=sin(pi x)sin(pi y)sin(pi z), FP64 CG + 7-level V(1,1) "
                                 "Chebyshev(1)/Jacobi, 5^3 Cholesky, rtol 1e-10 (preconditioned "
                                 "norm, PETSc's CG default)",
                     "grid": [257, 257, 257], "levels": 7, "its": rep.its,
                     "true_relres": rep.true_relres, "errinf": einf,
                     "errinf_closed_form": 1.254995e-05, "parallelism": "single GPU"}
    traffic = measured_traffic(capi.KERNEL_NAMES[dominant])
    out["roofline"] = {"kernel": capi.KERNEL_NAMES[dominant], "bound": "hbm",
                       "achieved": roof["gbs"] if roof else None, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": (roof["gbs"] / HBM_PEAK_GBS) if roof else None,
                       "traffic": traffic["traffic_bytes"] if traffic else None,
                       "traffic_kernel": traffic["kernel"] if traffic else None,
                       "traffic_source": TRAFFIC_FILE if traffic else None,
                       "avg_ms": roof["avg_ms"] if roof else None,
                       "algorithmic_bytes_per_launch": roof["bytes_per_launch"] if roof else None,
                       "launches": roof["launches"] if roof else None}
    out["kernel_ms_per_solve"] = breakdown
    # whole solve: algorithmic bytes of every launch (all levels) / solve time
    out["solve_algorithmic_gb"] = solve_bytes / 1e9
    out["solve_effective_gbs"] = solve_bytes / (ms * 1e-3) / 1e9
    out["stencil_apply_257"] = apply257

    # ---- C3: 513^3 mixed vs FP64 -----------------------------------------
    if not a.no_513:
        c3 = {}
        sols = {}
        for prec in ("double", "mixed"):
            extra = " -ksp_norm_type unpreconditioned" if prec == "double" else " -p4g_precision mixed"
            print("bench: 513^3 %s" % prec, file=sys.stderr, flush=True)
            with capi.Session(C3 + extra) as s:
                tot, l3, r3 = time_solves(s, 3, 1)
                m3 = tot / 3 * 1e3
                e3, _ = s.errors()
                sols[prec] = s.solution()
                c3[prec] = {"dof_per_s": s.n / (m3 * 1e-3), "ms_per_solve": m3, "its": r3[-1].its,
                            "inner_its": r3[-1].inner_its, "true_relres": r3[-1].true_relres,
                            "errinf_vs_exact": e3}
                if prec == "double":
                    c3["stencil_apply_513"] = stencil_apply_roofline(s)
        import numpy as np
        d = float(np.max(np.abs(sols["mixed"] - sols["double"])) / np.max(np.abs(sols["double"])))
        del sols
        c3["speedup_mixed_over_fp64"] = c3["mixed"]["dof_per_s"] / c3["double"]["dof_per_s"]
        c3["mixed_vs_fp64_rel_inf_diff"] = d
        c3["workload"] = ("3D Poisson 513^3 (135,005,697 DOF), f = manufactured + N(0,1e-3) noise, "
                          "8-level V(2,2) Chebyshev/Jacobi; FP64: CG on the FP64 hierarchy; mixed: "
                          "FP64 iterative refinement around FP32 PCG + FP32 V-cycle; both to "
                          "|b-Ax|_2/|b|_2 <= 1e-10")
        out["mixed_vs_fp64_513"] = c3

    # ---- C4: 257^3 variable coefficient, FP64 and mixed ---------------------
    if not a.no_513:
        c4 = {}
        for prec in ("double", "mixed"):
            extra = " -ksp_norm_type unpreconditioned" if prec == "double" else " -p4g_precision mixed"
            print("bench: 257^3 variable coefficient %s" % prec, file=sys.stderr, flush=True)
            with capi.Session(C4 + extra) as s:
                tot, l4, r4 = time_solves(s, 3, 1)
                m4 = tot / 3 * 1e3
                e4, _ = s.errors()
                c4[prec] = {"dof_per_s": s.n / (m4 * 1e-3), "ms_per_solve": m4, "its": r4[-1].its,
                            "true_relres": r4[-1].true_relres, "errinf_vs_exact": e4}
        c4["speedup_mixed_over_fp64"] = c4["mixed"]["dof_per_s"] / c4["double"]["dof_per_s"]
        c4["workload"] = ("3D -div(k grad u) = f on 257^3 vertices (256^3 cells), k = exp(G), G a seeded "
                          "Gaussian random field (random Fourier features, sigma 1, correlation length 0.1), "
                          "face coefficient = mean of the 4 cells on its edge, coarse k by averaging; "
                          "7-level V(1,1) Chebyshev/Jacobi; both to |b-Ax|_2/|b|_2 <= 1e-10")
        out["varcoef_257"] = c4

    # ---- C5: serving stream ------------------------------------------------
    if not a.no_serve:
        print("bench: serving stream", file=sys.stderr, flush=True)
        out["serving"] = serving()

    # ---- obstacle problem (c/ch12/obstacle.c), grid-sequenced to 2049^2 ---
    if not a.no_obstacle:
        print("bench: obstacle 2049^2", file=sys.stderr, flush=True)
        out["obstacle_2049"] = obstacle(not a.no_cpu)
        out["obstacle_cluster_2049"] = obstacle_cluster(not a.no_cpu)
    if not a.no_minimal:
        out["minimal_2049"] = minimal(not a.no_cpu)

    # ---- CPU baseline ------------------------------------------------------
    if not a.no_cpu:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(C2)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
