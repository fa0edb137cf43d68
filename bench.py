"""Benchmark: 3D Poisson DOF/s to rtol 1e-10 with CG + geometric multigrid on
one MI355X, mixed precision vs FP64, p50/p99 solve latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

A step is one KSP solve (from x = 0) of the 513^3 system (135,005,697 DOF =
mx*my*mz, the reference's global Vec size): f = manufactured sine right-hand
side + N(0, 1e-3) noise, CG + 8-level V(2,2) Chebyshev/Jacobi multigrid, to a
true FP64 relative residual |b - A x|_2 / |b|_2 <= 1e-10.  `value` is the FP64
MG-CG path; the mixed-precision path (FP32 V-cycle and Krylov vectors inside
FP64 iterative refinement) is timed the same way beside it.  F(u0) is already
resident in HBM; set-up (callbacks, hierarchy, eigen-estimates, coarse
factorization) is outside the timed region, as KSPSolve is timed apart from
SNESSolve in c/ch7/study/optimal3.sh:22-23.  Prints ONE JSON line.
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


TRAFFIC_FILES = {513: "profiles/hbm_traffic_513cubed.json",
                 257: "profiles/hbm_traffic_fp64_257cubed.json"}


def measured_traffic(grid, prec, cls_name):
    """HBM bytes per finest-level launch of a kernel class on the grid^3
    problem in precision prec ("double" or "mixed"), from the calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes recorded in TRAFFIC_FILES[grid]
    (None when absent)."""
    rel = TRAFFIC_FILES.get(grid)
    path = os.path.join(ROOT, rel) if rel else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {})
    if "classes" in d:   # per precision: class -> the finest-level kernel's name
        name = d["classes"].get(prec, {}).get(cls_name)
        t = k.get(prec, {}).get(name) if name else None
        if t is not None:
            t = dict(t, kernel=name)
    else:                # keyed by class
        t = k.get(cls_name)
    if t is not None:
        t = dict(t, source=rel)
    return t


def kernel_roofline(sess, cls):
    """Average duration and algorithmic bytes per launch of kernel class
    `cls` on the finest level, from HIP events recorded around each launch
    on the solver's stream during the solves since the last reset."""
    n, ms, by = sess.kernel_stats(cls, fine_only=True)
    if n == 0 or ms <= 0:
        return None
    return {"launches": n, "avg_ms": ms / n, "bytes_per_launch": by / n,
            "gbs": (by / n) / (ms / n * 1e-3) / 1e9}


def device_copy_rate(nbytes=1 << 30, reps=5):
    """The runtime's device-to-device copy of 1 GiB (hipMemcpyDtoD): moved
    bytes (read + write) per second, median of `reps` after a warm-up, host
    clock around each synchronous copy.  The practical HBM rate a streaming
    kernel can aim at, beside the 8 TB/s published peak."""
    from p4pdes_amd import capi
    L = capi.lib()
    src = L.p4g_device_malloc(nbytes)
    dst = L.p4g_device_malloc(nbytes)
    if not src or not dst:
        return None
    try:
        capi.check(L.p4g_memcpy_dtod(dst, src, nbytes))
        capi.check(L.p4g_device_synchronize())
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            capi.check(L.p4g_memcpy_dtod(dst, src, nbytes))
            capi.check(L.p4g_device_synchronize())
            ts.append(time.perf_counter() - t0)
    finally:
        L.p4g_device_free(src)
        L.p4g_device_free(dst)
    return 2.0 * nbytes / sorted(ts)[reps // 2] / 1e9


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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, label, nsolves=3, sample_its=2):
    """The oracle (C restatement, CSR operator as PETSc's AIJ), compiled here
    with -march=native, on the host cores given to this job: one warm-up,
    then nsolves KSP solves of the same system, median; then one solve on a
    single thread capped at sample_its iterations, scaled to the full
    iteration count (k iterations apply the preconditioner k + 1 times), as
    the 1-core figure.  Set-up is timed apart."""
    import tempfile
    from oracle import oracle_py
    nproc = os.cpu_count()
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    try:
        libpath = oracle_py.build_native(tempfile.mkdtemp(prefix="oracle_native_"))
        build = "gcc -O3 -march=native (built on this host)"
    except Exception as e:      # noqa: BLE001 - fall back to the shipped x86-64-v3 build
        libpath = None
        build = "gcc -O3 -march=x86-64-v3 (native build failed: %s)" % e
    r = oracle_py.fish_time_sample(args, threads, 1, nsolves, sample_its, libpath=libpath)
    _, mx, my, mz = oracle_py.grid_shape(args)
    n = mx * my * mz
    times = r["times_ms"]
    med = sorted(times)[len(times) // 2]
    its = r["res"].its
    # a PCG solve of k iterations applies the preconditioner k + 1 times
    one_solve_ms = r["sample_ms"] / (r["sample_its"] + 1) * (its + 1)
    return {"value": n / (med * 1e-3), "unit": "DOF/s", "cores": threads, "kind": "port",
            "sample": "%d KSP solves of the full %s system after one warm-up (median %.2f s, total %.1f s; "
                      "set-up %.1f s apart), oracle CSR CG + multigrid, %d OpenMP threads, %d its"
                      % (nsolves, label, med * 1e-3, sum(times) * 1e-3, r["setup_ms"] * 1e-3, threads, its),
            "median_solve_s": med * 1e-3, "its": its,
            "one_core": {"value": n / (one_solve_ms * 1e-3), "unit": "DOF/s", "cores": 1,
                         "sample": "one solve capped at %d iterations on 1 thread (%.2f s), scaled by "
                                   "preconditioner applications to the %d-iteration solve"
                                   % (r["sample_its"], r["sample_ms"] * 1e-3, its)},
            "host": {"nproc": nproc, "affinity_cpus": affinity, "omp_num_threads": omp,
                     "model": _cpu_model(), "oracle_build": build}}


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


def serving(rate=40.0, nreq=80, seed=2026, budget=64 << 30, server_opts=""):
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
    alone = {}
    with capi.Server(budget, server_opts) as srv:
        # warm-up burst, not timed: two requests of every size at once, so
        # the server's vector and staging pools hold what the stream needs
        # (a cold pool allocates inside the first requests' latency)
        warm = []
        for dim, r in SERVE_SHAPES:
            m = 2 ** (r + 1) + 1
            bw, _ = sine_mode_request(dim, m, seed * 13 + r)
            warm += [srv.submit(sess[(dim, r)].solver, bw) for _ in range(2)]
        for w in warm:
            srv.wait(w)
        st0 = srv.stats()
        t0 = time.perf_counter()
        ids = []
        for q, (dim, r, m, b, _) in enumerate(reqs):
            delay = arrivals[q] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(sess[(dim, r)].solver, b))
        for q, rid in enumerate(ids):
            out.append(srv.wait(rid))
        wall = time.perf_counter() - t0
        st = srv.stats()
        # each size alone through the same server: three requests one at a time
        for dim, r in SERVE_SHAPES:
            m = 2 ** (r + 1) + 1
            b, _ = sine_mode_request(dim, m, seed * 7 + r)
            lat_alone = []
            for _ in range(3):
                lat_alone.append(srv.wait(srv.submit(sess[(dim, r)].solver, b))[1].solve_ms)
            alone[(dim, r)] = sorted(lat_alone)[1]
    errs = []
    for q, (dim, r, m, b, xh) in enumerate(reqs):
        if xh is not None:
            errs.append(float(np.max(np.abs(out[q][0] - xh)) / np.max(np.abs(xh))))
    for s in sess.values():
        s.close()
    its = [rep.its for _, rep in out]
    per_size = {}
    for (dim, r) in SERVE_SHAPES:
        lat = [out[q][1].solve_ms for q in range(len(reqs)) if (reqs[q][0], reqs[q][1]) == (dim, r)]
        m = 2 ** (r + 1) + 1
        per_size["%d^%d" % (m, dim)] = {"requests": len(lat), "alone_ms": alone[(dim, r)],
                                         "mix_p50_ms": pct(lat, 0.5), "mix_p99_ms": pct(lat, 0.99)}
    largest_alone = max(alone.values())
    slow = sorted(range(len(reqs)), key=lambda q: -out[q][1].solve_ms)[:5]
    slowest = [{"size": "%d^%d" % (reqs[q][2], reqs[q][0]), "latency_ms": out[q][1].solve_ms,
                "queue_ms": out[q][1].setup_ms, "its": out[q][1].its, "arrival_s": float(arrivals[q])}
               for q in slow]
    lat_all = [rep.solve_ms for _, rep in out]
    p99 = pct(lat_all, 0.99)
    dof = sum(float(reqs[q][2] ** reqs[q][0]) for q in range(len(reqs)))
    return {"p50_ms": pct(lat_all, 0.5), "p99_ms": p99, "mean_ms": float(np.mean(lat_all)),
            "max_ms": max(lat_all), "largest_alone_ms": largest_alone, "p99_over_largest_alone": p99 / largest_alone,
            "per_size": per_size, "slowest": slowest,
            "queue_p50_ms": pct([rep.setup_ms for _, rep in out], 0.5), "throughput_dof_per_s": dof / wall,
            "requests": len(out), "rate_per_s": rate, "wall_s": wall,
            "rounds": int(st.rounds - st0.rounds), "graph_captures": int(st.graph_captures - st0.graph_captures),
            "peak_hbm_gb": st.peak_bytes / 2 ** 30, "its_median": float(np.median(its)),
            "max_rel_err_vs_exact_discrete": max(errs) if errs else None,
            "all_converged": all(rep.reason > 0 for _, rep in out),
            "workload": "Poisson arrivals (seeded) at %.0f/s of %d solves, sizes uniform over 33^3, 65^3, "
                        "129^3, 257^3, 257^2, 1025^2, 2049^2; RHS = 8 random sine modes, N(0,1) "
                        "amplitudes; zero Dirichlet; mixed precision to rtol 1e-8; V(1,1) "
                        "Chebyshev/Jacobi to 5^d; HBM budget %d GiB; after an untimed warm-up burst of two "
                        "requests per size (the server's pools filled); latency = submit -> result in the "
                        "caller's array" % (rate, nreq, budget >> 30)}


COPY_BYTES = 1 << 30
C3_FP64 = C3 + " -ksp_norm_type unpreconditioned"
C3_MIXED = C3 + " -p4g_precision mixed"


def profile_workload(which):
    """Workload for hardware-counter and trace passes: two solves of the
    513^3 system in FP64 and two in mixed precision ("513"), or two 257^3
    FP64 solves ("257"); then a device copy of COPY_BYTES (read COPY_BYTES,
    write COPY_BYTES) whose counter readings calibrate the counters' bytes."""
    from p4pdes_amd import capi
    cases = [C2] if which == "257" else [C3_FP64, C3_MIXED]
    for opts in cases:
        print("profile workload: %s" % opts, flush=True)
        with capi.Session(opts) as s:
            for _ in range(2):
                s.solve()
            capi.check(capi.lib().p4g_device_synchronize())
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


def timed_config(opts, steps, warmup, grid, prec, roofline=True):
    """Set up one fish session and time `steps` solves after `warmup`; with
    roofline, first time every kernel class once (one untimed solve) to find
    the dominant class, then event-time its finest-level launches during the
    timed solves (one event pair per launch on the solver's stream)."""
    from p4pdes_amd import capi
    out = {}
    with capi.Session(opts) as s:
        s.solve()
        dominant = None
        if roofline:
            s.reset_stats()
            for c in range(capi.K_COUNT):
                s.set_timing(c, True)
            s.solve()
            by_cls = {c: s.kernel_stats(c)[1] for c in range(capi.K_COUNT)}
            solve_bytes = sum(s.kernel_stats(c)[2] for c in range(capi.K_COUNT))
            for c in range(capi.K_COUNT):
                s.set_timing(c, False)
            dominant = max((c for c in by_cls if c not in (capi.K_OTHER, capi.K_REDUCE)),
                           key=lambda c: by_cls[c])
            out["kernel_ms_per_solve_evented"] = {capi.KERNEL_NAMES[c]: round(by_cls[c], 4) for c in by_cls}
            out["solve_algorithmic_gb"] = solve_bytes / 1e9
            s.reset_stats()
            s.set_timing(dominant, True, fine_only=True)
        total, lat, reps = time_solves(s, steps, warmup)
        if roofline:
            s.set_timing(dominant, False)
            roof = kernel_roofline(s, dominant)
            name = capi.KERNEL_NAMES[dominant]
            traffic = measured_traffic(grid, prec, name)
            out["roofline"] = {"kernel": name, "bound": "hbm",
                               "achieved": roof["gbs"] if roof else None, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": (roof["gbs"] / HBM_PEAK_GBS) if roof else None,
                               "traffic": traffic["traffic_bytes"] if traffic else None,
                               "traffic_kernel": traffic["kernel"] if traffic else None,
                               "traffic_source": traffic["source"] if traffic else None,
                               "avg_ms": roof["avg_ms"] if roof else None,
                               "algorithmic_bytes_per_launch": roof["bytes_per_launch"] if roof else None,
                               "launches": roof["launches"] if roof else None}
        einf, _ = s.errors()
        ms = total / steps * 1e3
        rep = reps[-1]
        out.update({"dof_per_s": s.n / (ms * 1e-3), "ms_per_solve": ms,
                    "latency_ms": {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99)},
                    "its": rep.its, "inner_its": rep.inner_its, "true_relres": rep.true_relres,
                    "errinf_vs_exact": einf, "n": s.n})
        if roofline:
            out["solve_effective_gbs"] = out["solve_algorithmic_gb"] / (ms * 1e-3)
        if prec == "double":
            out["stencil_apply"] = stencil_apply_roofline(s)
        out["_solution"] = s.solution() if grid == 513 else None
    return out


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


FISH_CLUSTER = "-fsh_dim 3 -da_refine 7 -pc_mg_levels 6 -pc_type mg"


def fish_cluster(with_oracle):
    """fish as c/ch8/cluster.sh:58-63 runs it: 257^3 (manuexp), CG + 6-level
    multigrid with PETSc's default level smoother, Chebyshev(2) with SOR (the
    sweeps as pipelined strips, one launch per direction), rtol 1e-5.  KSP
    solve time (median of 3 after a warm-up), iterations and the error line;
    the same solve with one launch per hyperplane; the oracle's solve time."""
    from p4pdes_amd import capi

    def run(opts, n):
        with capi.Session(opts) as s:
            s.solve()
            capi.check(capi.lib().p4g_device_synchronize())
            ts = []
            for _ in range(n):
                t0 = time.perf_counter()
                rep = s.solve()
                capi.check(capi.lib().p4g_device_synchronize())
                ts.append((time.perf_counter() - t0) * 1e3)
            return sorted(ts)[len(ts) // 2], rep, s.n

    ms, rep, n = run(FISH_CLUSTER, 3)
    ms_planes, rep_p, _ = run(FISH_CLUSTER + " -p4g_sor_pipeline 0", 1)
    o = {"workload": "fish -fsh_dim 3 -da_refine 7 -pc_mg_levels 6 -pc_type mg (c/ch8/cluster.sh:63): 257^3, "
                     "CG + V(2,2) Chebyshev/SOR multigrid, 9^3 coarse LU, rtol 1e-5",
         "options": FISH_CLUSTER, "ms_per_solve": ms, "dof_per_s": n / (ms * 1e-3), "its": rep.its,
         "ms_per_solve_hyperplane_sor": ms_planes, "its_hyperplane_sor": rep_p.its}
    if with_oracle:
        from oracle import oracle_py
        nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        times, ores = oracle_py.fish_time(FISH_CLUSTER, nthreads=nt, warmup=0, nsolves=1)
        o["oracle_ms_per_solve"] = times[0]
        o["oracle_its"] = ores.its
        o["oracle_threads"] = nt
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
    ap.add_argument("--no-extra", action="store_true",
                    help="only the 513^3 headline (no 257^3, variable-coefficient, serving, "
                         "obstacle or minimal-surface legs)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-obstacle", action="store_true", help="skip the obstacle (VI) solve")
    ap.add_argument("--no-minimal", action="store_true", help="skip the minimal-surface solve")
    ap.add_argument("--profile-workload", choices=("513", "257"), default=None,
                    help="only run the counter-collection workload: two solves of the 513^3 system "
                         "in each precision (or two 257^3 FP64 solves) and a 1 GiB device copy (the "
                         "calibration copy of known size); prints no JSON")
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
        profile_workload(a.profile_workload)
        return
    import numpy as np

    out = {"metric": METRIC, "unit": "DOF/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (manufactured sine right-hand side plus seeded N(0,1e-3) noise; "
                   "closed-form noise-free solution)"}

    # ---- headline: C3 513^3, FP64 MG-CG (value) and mixed precision beside it
    print("bench: 513^3 FP64, %d + %d solves" % (a.warmup, a.steps), file=sys.stderr, flush=True)
    d = timed_config(C3_FP64, a.steps, a.warmup, 513, "double")
    print("bench: 513^3 mixed, %d + %d solves" % (a.warmup, a.steps), file=sys.stderr, flush=True)
    m = timed_config(C3_MIXED, a.steps, a.warmup, 513, "mixed")
    diff = float(np.max(np.abs(m["_solution"] - d["_solution"])) / np.max(np.abs(d["_solution"])))
    del d["_solution"], m["_solution"]
    out["value"] = d["dof_per_s"]
    out["ms_per_step"] = d["ms_per_solve"]
    out["latency_ms"] = d["latency_ms"]
    out["config"] = {"workload": "C3: 3D Poisson 513^3 vertex grid (135,005,697 DOF), 7-point stencil, "
                                 "f = 3 pi^2 sin(pi x)sin(pi y)sin(pi z) + N(0,1e-3) noise (Philox, seed 0), "
                                 "CG + 8-level V(2,2) Chebyshev(2)/Jacobi, 5^3 Cholesky, to "
                                 "|b-Ax|_2/|b|_2 <= 1e-10; value = FP64 path, mixed = FP32 V-cycle in "
                                 "FP64 iterative refinement",
                     "grid": [513, 513, 513], "levels": 8, "its": d["its"],
                     "true_relres": d["true_relres"], "errinf": d["errinf_vs_exact"],
                     "errinf_closed_form_noise_free": 3.137469e-06, "parallelism": "single GPU"}
    out["roofline"] = d.pop("roofline")
    copy_gbs = device_copy_rate()
    if out["roofline"] and copy_gbs:
        # the practical streaming rate of this GPU beside the published peak
        out["roofline"]["device_copy_gbs"] = copy_gbs
        out["roofline"]["frac_of_device_copy"] = out["roofline"]["achieved"] / copy_gbs
    out["mixed"] = {"value": m["dof_per_s"], "ms_per_step": m["ms_per_solve"], "latency_ms": m["latency_ms"],
                    "its": m["its"], "inner_its": m["inner_its"], "true_relres": m["true_relres"],
                    "errinf": m["errinf_vs_exact"], "dtype": "f32 V-cycle and Krylov vectors, f64 iterate, "
                    "residual replacement, dots and coarse solve",
                    "speedup_over_fp64": m["dof_per_s"] / d["dof_per_s"],
                    "rel_inf_diff_vs_fp64": diff, "roofline": m.get("roofline"),
                    "kernel_ms_per_solve_evented": m.get("kernel_ms_per_solve_evented"),
                    "solve_algorithmic_gb": m.get("solve_algorithmic_gb")}
    out["fp64_detail"] = {"kernel_ms_per_solve_evented": d.get("kernel_ms_per_solve_evented"),
                          "solve_algorithmic_gb": d.get("solve_algorithmic_gb"),
                          "solve_effective_gbs": d.get("solve_effective_gbs"),
                          "stencil_apply_513": d.get("stencil_apply")}

    if not a.no_extra:
        # ---- C2: 257^3 FP64 -------------------------------------------------
        print("bench: 257^3 FP64", file=sys.stderr, flush=True)
        c2 = timed_config(C2, a.steps, a.warmup, 257, "double")
        del c2["_solution"]
        c2["errinf_closed_form"] = 1.254995e-05
        c2["workload"] = ("C2: 3D Poisson 257^3 (16,974,593 DOF), u = sin(pi x)sin(pi y)sin(pi z), FP64 CG + "
                          "7-level V(1,1) Chebyshev(1)/Jacobi, 5^3 Cholesky, rtol 1e-10 (preconditioned norm, "
                          "PETSc's CG default)")
        out["fp64_257"] = c2

        # ---- C4: 257^3 variable coefficient, FP64 and mixed -----------------
        c4 = {}
        for prec in ("double", "mixed"):
            extra = " -ksp_norm_type unpreconditioned" if prec == "double" else " -p4g_precision mixed"
            print("bench: 257^3 variable coefficient %s" % prec, file=sys.stderr, flush=True)
            r4 = timed_config(C4 + extra, 3, 1, 257, prec, roofline=False)
            r4.pop("_solution", None)
            r4.pop("stencil_apply", None)
            c4[prec] = r4
        c4["speedup_mixed_over_fp64"] = c4["mixed"]["dof_per_s"] / c4["double"]["dof_per_s"]
        c4["workload"] = ("3D -div(k grad u) = f on 257^3 vertices (256^3 cells), k = exp(G), G a seeded "
                          "Gaussian random field (random Fourier features, sigma 1, correlation length 0.1), "
                          "face coefficient = mean of the 4 cells on its edge, coarse k by averaging; "
                          "7-level V(1,1) Chebyshev/Jacobi; both to |b-Ax|_2/|b|_2 <= 1e-10")
        out["varcoef_257"] = c4

        # ---- C5: serving stream ----------------------------------------------
        if not a.no_serve:
            print("bench: serving stream", file=sys.stderr, flush=True)
            out["serving"] = serving()

        # ---- obstacle problem (c/ch12/obstacle.c), grid-sequenced to 2049^2
        if not a.no_obstacle:
            print("bench: obstacle 2049^2", file=sys.stderr, flush=True)
            out["obstacle_2049"] = obstacle(not a.no_cpu)
            out["obstacle_cluster_2049"] = obstacle_cluster(not a.no_cpu)
        if not a.no_minimal:
            print("bench: minimal surface 2049^2", file=sys.stderr, flush=True)
            out["minimal_2049"] = minimal(not a.no_cpu)
        print("bench: fish 257^3 with Chebyshev/SOR (cluster.sh)", file=sys.stderr, flush=True)
        out["fish_cluster_257"] = fish_cluster(not a.no_cpu)

    # ---- CPU baseline: the same 513^3 FP64 system -----------------------------
    if not a.no_cpu:
        print("bench: CPU baseline (oracle, 513^3)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(C3_FP64, "513^3")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
