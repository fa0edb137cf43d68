// Reduced-space Newton for variational inequalities on the device: PETSc's
// SNESSolve_VINEWTONRSLS with SNESLineSearchApply_BT (alpha 0) or _Basic and
// SNESConvergedDefault, as c/ch12/obstacle.c:118-121 selects it.  The
// scalar decisions (convergence tests, the line-search model) run on the
// host; every vector pass runs on the device.  A Newton step solves the
// inactive-set system in full-size vectors with the active vertices as
// null rows (k_vi_nullrow27); under -pc_type mg the coarse operators are
// the Galerkin products of that operator, which is what PETSc's
// PCMGSetGalerkin under RSLS computes on its coarse inactive sets.
// Included once, at the end of p4g_solver.hip.

namespace p4g {

struct ViRun {
    p4g_solver_s* s = nullptr;
    ViGrid g;
    long n = 0;
    int nb = 1;
    hipStream_t st = nullptr;
    double *F0 = nullptr, *X = nullptr, *F = nullptr, *Y = nullptr, *W = nullptr, *G = nullptr,
           *Fm = nullptr, *xl = nullptr, *xu = nullptr;
    double* S = nullptr;   // scalar block: the second half of cx.S
    // -snes_type vinewtonssls: the merit function is |phi| (phi the Fischer-
    // Burmeister function of X), a trial point's phi goes to G and its
    // residual to FW
    bool ssls = false;
    double *phi = nullptr, *FW = nullptr;

    int sums(double& a, double& b) {
        P4G_TRY(s->cx.finalize(nb, S, FIN_RAW, 0));
        P4G_TRY(s->cx.readS(s->cx.hS));
        a = s->cx.hS[16 + S_TMP0];
        b = s->cx.hS[16 + S_TMP1];
        return 0;
    }
    // f = A x + F0 (rhs) or A x
    int fun(const double* x, double* f, bool rhs) {
        k_vi_fun<<<nb, 256, 0, st>>>(g, x, rhs ? F0 : nullptr, f);
        LAUNCH_CHECK();
        return 0;
    }
    // |f| on the inactive set of (x, f), and |x|
    int norms(const double* x, const double* f, double& fn, double& xn) {
        k_vi_sets<<<nb, 256, 0, st>>>(n, x, f, xl, xu, nullptr, nullptr, s->cx.part);
        LAUNCH_CHECK();
        double a, b;
        P4G_TRY(sums(a, b));
        fn = std::sqrt(a);
        xn = std::sqrt(b);
        return 0;
    }
    // (x, y) and |y|
    int dot(const double* x, const double* y, double& xy, double& yn) {
        k_vi_dot<<<nb, 256, 0, st>>>(n, x, y, s->cx.part);
        LAUNCH_CHECK();
        double a, b;
        P4G_TRY(sums(a, b));
        xy = a;
        yn = std::sqrt(b);
        return 0;
    }
    // w = proj(x0 - lambda y), gv = F(w), gn = |gv| on w's inactive set;
    // SSLS: FW = F(w), gv = phi(w), gn = |gv|
    int trial(const double* x0, const double* y, double lambda, double* w, double* gv, double& gn) {
        k_vi_trial<<<nb, 256, 0, st>>>(n, x0, y, lambda, xl, xu, w);
        LAUNCH_CHECK();
        if (ssls) {
            P4G_TRY(fun(w, FW, true));
            k_ssls_phi<<<nb, 256, 0, st>>>(n, w, xl, FW, gv);
            LAUNCH_CHECK();
            double t;
            return dot(gv, gv, t, gn);
        }
        P4G_TRY(fun(w, gv, true));
        double xn;
        return norms(w, gv, gn, xn);
    }
};

static int asm_setup(p4g_solver_s* s, ViRun& R);

// the finest operator of the current active set and, under -pc_type mg, its
// Galerkin coarse levels, coarse factor and Chebyshev bounds
static int vi_operator(p4g_solver_s* s, ViRun& R) {
    const int L = s->nlev - 1;
    Engine<double>& e = *s->e64;
    // per-level coefficient tables exist only under -pc_type mg -pc_mg_galerkin
    if ((int)s->g27.size() != s->nlev) s->g27.assign(s->nlev, nullptr);
    if ((int)e.g27.size() != s->nlev) e.g27.assign(s->nlev, nullptr);
    k_vi_nullrow27<<<R.nb, 256, 0, R.st>>>(R.g, s->vi_act, s->vi_g27);
    LAUNCH_CHECK();
    s->lv[L].gal = true;
    s->g27[L] = s->vi_g27;
    e.g27[L] = s->vi_g27;
    e.vidiag = s->vi_g27 + 13 * R.n;
    if (s->pctype == PCT_ILU) return ilu27_pivots(s->lv[L], s->vi_g27, s->ilu_u[L], R.st);   // PETSc's default PC
    if (s->pctype == PCT_ASM) return asm_setup(s, R);
    if (s->pctype != PCT_MG) return 0;
    P4G_TRY(galerkin_levels(s));
    for (int l = 0; l < L; ++l) e.g27[l] = s->g27[l];
    for (int l = 0; l <= L; ++l)   // the pipelined SOR sweeps read the operator as packed records
        if (s->lv[l].pipe)
            k_sor9_pack<<<vblocks(s->lv[l].n), 256, 0, R.st>>>(s->lv[l].n, s->g27[l], s->lv[l].pipe->pk);
    LAUNCH_CHECK();
    if (L > 0) {
        std::vector<double> Ainv;
        P4G_TRY(coarse_factor(s->lv[0], s->g27[0], Ainv));
        P4G_TRY(e.set_coarse_inverse(Ainv));
    }
    if (s->smoother == SM_CHEB)   // one level: the level smoother is the preconditioner
        for (int l = L > 0 ? 1 : 0; l <= L; ++l) P4G_TRY(estimate_level(s, l));
    return 0;
}

// dense inverse by Gauss-Jordan elimination with partial pivoting (the
// subdomain solve is exact up to rounding, as PETSc's subdomain LU is)
static bool dense_inverse(int m, std::vector<double>& A, double* X) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) X[(long)i * m + j] = i == j ? 1.0 : 0.0;
    for (int c = 0; c < m; ++c) {
        int pr = c;
        for (int r = c + 1; r < m; ++r)
            if (std::fabs(A[(long)r * m + c]) > std::fabs(A[(long)pr * m + c])) pr = r;
        if (A[(long)pr * m + c] == 0.0) return false;
        if (pr != c)
            for (int j = 0; j < m; ++j) {
                std::swap(A[(long)c * m + j], A[(long)pr * m + j]);
                std::swap(X[(long)c * m + j], X[(long)pr * m + j]);
            }
        const double piv = A[(long)c * m + c];
        for (int j = 0; j < m; ++j) { A[(long)c * m + j] /= piv; X[(long)c * m + j] /= piv; }
        for (int r = 0; r < m; ++r) {
            if (r == c) continue;
            const double f = A[(long)r * m + c];
            if (f == 0.0) continue;
            for (int j = 0; j < m; ++j) {
                A[(long)r * m + j] -= f * A[(long)c * m + j];
                X[(long)r * m + j] -= f * X[(long)c * m + j];
            }
        }
    }
    return true;
}

// PCASM over the reduced (inactive-set) operator, as PETSc's default PCASM
// runs on -p4g_ranks processes: one subdomain per process, its inactive
// vertices of the DMDA's ownership block (DMSetUp_DA_2D's process grid),
// grown by -pc_asm_overlap levels of the reduced operator's graph
// (MatIncreaseOverlap), solved exactly (-sub_pc_type lu), the restricted
// form (a subdomain's solution kept on its own rows only).  The subdomains
// follow the active set, so they are rebuilt on the host every Newton step
// and uploaded as index lists and dense inverses (subdomains up to 2048
// vertices).
static int asm_setup(p4g_solver_s* s, ViRun& R) {
    const int mx = R.g.mx, my = R.g.my;
    const long n = R.n;
    if (R.g.dim != 2) return set_error(PETSC_ERR_SUP, "-pc_type asm is implemented for 2D grids");
    std::vector<unsigned char> act(n);
    HIP_TRY(hipStreamSynchronize(R.st));
    HIP_TRY(hipMemcpy(act.data(), s->vi_act, n, hipMemcpyDeviceToHost));
    const int P = s->ranks;
    int m = (int)(0.5 + std::sqrt((double)mx * P / my));
    if (!m) m = 1;
    int nn = 1;
    while (m > 0) {
        nn = P / m;
        if (m * nn == P) break;
        --m;
    }
    if (mx > my && m < nn) std::swap(m, nn);
    std::vector<int> xs(m + 1, 0), ys(nn + 1, 0);
    for (int q = 0; q < m; ++q) xs[q + 1] = xs[q] + mx / m + (mx % m > q);
    for (int q = 0; q < nn; ++q) ys[q + 1] = ys[q] + my / nn + (my % nn > q);
    std::vector<int> rx(mx), ry(my);
    for (int i = 0, q = 0; i < mx; ++i) { while (i >= xs[q + 1]) ++q; rx[i] = q; }
    for (int j = 0, q = 0; j < my; ++j) { while (j >= ys[q + 1]) ++q; ry[j] = q; }
    auto inner = [&](long v) {
        const int i = (int)(v % mx), j = (int)(v / mx);
        return i > 0 && i < mx - 1 && j > 0 && j < my - 1;
    };
    const long nb4[4] = {-(long)mx, -1, 1, (long)mx};
    std::vector<int> idx_all;
    std::vector<double> inv_all;
    std::vector<long> rows;
    std::vector<char> in(n), nx(n);
    std::vector<long> loc(n, -1);
    for (int q = 0; q < P; ++q) {
        for (long v = 0; v < n; ++v)
            in[v] = !act[v] && (ry[v / mx] * m + rx[v % mx]) == q;
        for (int o = 0; o < s->asm_overlap; ++o) {
            nx = in;
            for (long v = 0; v < n; ++v)
                if (in[v] && inner(v))
                    for (long d : nb4)
                        if (inner(v + d) && !act[v + d]) nx[v + d] = 1;
            in.swap(nx);
        }
        std::vector<long> list;
        for (long v = 0; v < n; ++v)
            if (in[v]) list.push_back(v);
        const long c = (long)list.size();
        if (c == 0) continue;
        if (c > MAX_COARSE)
            return set_error(PETSC_ERR_SUP, "-pc_type asm: a subdomain of %ld vertices (dense solves up to %d)", c,
                             MAX_COARSE);
        for (long k = 0; k < c; ++k) loc[list[k]] = k;
        std::vector<double> A((size_t)c * c, 0.0);
        for (long k = 0; k < c; ++k) {
            const long v = list[k];
            const bool iv = inner(v);
            A[(size_t)k * c + k] = iv ? R.g.d : R.g.db;
            if (!iv) continue;
            for (long d : nb4) {
                const long u = v + d;
                if (!inner(u) || act[u] || loc[u] < 0) continue;
                A[(size_t)k * c + loc[u]] = (d == 1 || d == -1) ? -R.g.sx : -R.g.sy;
            }
        }
        const long ioff = (long)inv_all.size(), xoff = (long)idx_all.size();
        inv_all.resize(inv_all.size() + (size_t)c * c);
        if (!dense_inverse((int)c, A, inv_all.data() + ioff))
            return set_error(PETSC_ERR_LIB, "-pc_type asm: singular subdomain matrix");
        for (long k = 0; k < c; ++k) {
            idx_all.push_back((int)list[k]);
            const long v = list[k];
            if ((ry[v / mx] * m + rx[v % mx]) == q) {   // owned: the restricted form writes it back
                rows.push_back(ioff + k * c);
                rows.push_back(xoff);
                rows.push_back(c);
                rows.push_back(v);
            }
        }
        for (long k = 0; k < c; ++k) loc[list[k]] = -1;
    }
    auto grow = [&](void** p, size_t* cap, size_t bytes) -> int {
        if (*cap >= bytes && *p) return 0;
        if (*p) hipFree(*p);
        *p = nullptr;
        *cap = 0;
        const size_t want = std::max<size_t>(bytes * 2, 4096);
        if (hipMalloc(p, want) != hipSuccess) return set_error(PETSC_ERR_MEM, "hipMalloc (PCASM)");
        *cap = want;
        return 0;
    };
    P4G_TRY(grow((void**)&s->asm_idx, &s->asm_cap_idx, idx_all.size() * sizeof(int)));
    P4G_TRY(grow((void**)&s->asm_inv, &s->asm_cap_inv, inv_all.size() * sizeof(double)));
    P4G_TRY(grow((void**)&s->asm_row, &s->asm_cap_row, rows.size() * sizeof(long)));
    if (!idx_all.empty()) HIP_TRY(hipMemcpy(s->asm_idx, idx_all.data(), idx_all.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!inv_all.empty()) HIP_TRY(hipMemcpy(s->asm_inv, inv_all.data(), inv_all.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!rows.empty()) HIP_TRY(hipMemcpy(s->asm_row, rows.data(), rows.size() * sizeof(long), hipMemcpyHostToDevice));
    s->asm_nrows = (long)rows.size() / 4;
    return 0;
}

// the semismooth Jacobian diag(Db) J + diag(Da) at the current X (SSLS) as
// the finest level's 27-point operator, and its ILU(0) pivots
static int ssls_operator(p4g_solver_s* s, ViRun& R) {
    const int L = s->nlev - 1;
    Engine<double>& e = *s->e64;
    if ((int)s->g27.size() != s->nlev) s->g27.assign(s->nlev, nullptr);
    if ((int)e.g27.size() != s->nlev) e.g27.assign(s->nlev, nullptr);
    k_ssls_op27<<<R.nb, 256, 0, R.st>>>(R.g, R.X, R.xl, R.F, s->vi_g27);
    LAUNCH_CHECK();
    s->lv[L].gal = true;
    s->g27[L] = s->vi_g27;
    e.g27[L] = s->vi_g27;
    e.vidiag = nullptr;
    if (s->pctype == PCT_ILU) return ilu27_pivots(s->lv[L], s->vi_g27, s->ilu_u[L], R.st);
    return 0;
}

// SNESConvergedDefault (SNESConvergedReason numbering)
static int vi_converged(const p4g_solver_s* s, int it, double xnorm, double snorm, double fnorm, double& ttol,
                        double& rnorm0) {
    if (it == 0) { ttol = fnorm * s->snes_rtol; rnorm0 = fnorm; }
    if (std::isnan(fnorm) || std::isinf(fnorm)) return -4;   // DIVERGED_FNORM_NAN
    if (fnorm < s->snes_atol) return 2;                      // CONVERGED_FNORM_ABS
    if (it) {
        if (fnorm <= ttol) return 3;                         // CONVERGED_FNORM_RELATIVE
        if (snorm < s->snes_stol * xnorm) return 4;          // CONVERGED_SNORM_RELATIVE
        if (s->snes_divtol > 0 && fnorm > s->snes_divtol * rnorm0) return -9;   // DIVERGED_DTOL
    }
    return 0;
}

// SNESLineSearchApply_BT with alpha = 0 (or _Basic).  On return X, F hold
// the accepted point (pointers swapped with the trial buffers), fnorm its
// norm, ynorm = lambda |Y|; returns 1 when the search failed.  SSLS: the
// function is phi, its slope (phi, Jss Y) with the semismooth Jacobian, and
// X, phi, F move to the accepted point.  The trial points are projected
// onto the bounds in both methods.
static int vi_line_search(p4g_solver_s* s, ViRun& R, double& fnorm, double& xnorm, double& ynorm, int* fail) {
    const double alpha = 0.0, steptol = 1e-12;
    double lambda = 1.0, ydot;
    *fail = 0;
    P4G_TRY(R.dot(R.X, R.X, ydot, xnorm));
    P4G_TRY(R.dot(R.Y, R.Y, ydot, ynorm));
    double gnorm = 0.0;
    if (s->ls_type == "basic") {
        P4G_TRY(R.trial(R.X, R.Y, 1.0, R.W, R.G, gnorm));
    } else {
        if (ynorm == 0.0) { *fail = 1; return 0; }
        // initial slope of |F|^2/2 along -Y: -(F, J Y)
        double initslope, wn;
        if (R.ssls) {
            P4G_TRY(s->e64->mat_mult(R.Y, R.W));
            P4G_TRY(R.dot(R.phi, R.W, initslope, wn));
        } else {
            P4G_TRY(R.fun(R.Y, R.W, false));
            P4G_TRY(R.dot(R.F, R.W, initslope, wn));
        }
        if (initslope > 0.0) initslope = -initslope;
        if (initslope == 0.0) initslope = -1.0;
        const double f = fnorm * fnorm;
        P4G_TRY(R.trial(R.X, R.Y, lambda, R.W, R.G, gnorm));
        double g = gnorm * gnorm;
        bool ok = 0.5 * g <= 0.5 * f + lambda * alpha * initslope;
        if (!ok) {
            // rellength = max |Y_i| / max(|X_i|, 1), on the host
            std::vector<double> hx(R.n), hy(R.n);
            HIP_TRY(hipStreamSynchronize(R.st));
            HIP_TRY(hipMemcpy(hx.data(), R.X, R.n * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hy.data(), R.Y, R.n * sizeof(double), hipMemcpyDeviceToHost));
            double rellength = 0.0;
            for (long q = 0; q < R.n; ++q)
                rellength = std::max(rellength, std::fabs(hy[q]) / std::max(std::fabs(hx[q]), 1.0));
            const double minlambda = steptol / rellength;
            // quadratic model
            double lt = -initslope / (g - f - 2.0 * lambda * initslope);
            double lprev = lambda, gprev = g;
            if (lt > 0.5 * lambda) lt = 0.5 * lambda;
            lambda = lt <= 0.1 * lambda ? 0.1 * lambda : lt;
            P4G_TRY(R.trial(R.X, R.Y, lambda, R.W, R.G, gnorm));
            g = gnorm * gnorm;
            ok = 0.5 * g < 0.5 * f + lambda * alpha * initslope;
            // cubic model
            while (!ok) {
                if (lambda <= minlambda) break;
                const double t1 = 0.5 * (g - f) - lambda * initslope;
                const double t2 = 0.5 * (gprev - f) - lprev * initslope;
                const double a = (t1 / (lambda * lambda) - t2 / (lprev * lprev)) / (lambda - lprev);
                const double b = (-lprev * t1 / (lambda * lambda) + lambda * t2 / (lprev * lprev)) / (lambda - lprev);
                double d = b * b - 3.0 * a * initslope;
                if (d < 0.0) d = 0.0;
                lt = (a == 0.0) ? -initslope / (2.0 * b) : (-b + std::sqrt(d)) / (3.0 * a);
                lprev = lambda;
                gprev = g;
                if (lt > 0.5 * lambda) lt = 0.5 * lambda;
                lambda = lt <= 0.1 * lambda ? 0.1 * lambda : lt;
                P4G_TRY(R.trial(R.X, R.Y, lambda, R.W, R.G, gnorm));
                g = gnorm * gnorm;
                ok = 0.5 * g <= 0.5 * f + lambda * alpha * initslope;
            }
        }
        *fail = ok ? 0 : 1;
    }
    std::swap(R.X, R.W);
    if (R.ssls) {
        std::swap(R.phi, R.G);
        std::swap(R.F, R.FW);
    } else {
        std::swap(R.F, R.G);
    }
    fnorm = gnorm;
    double xx;
    P4G_TRY(R.dot(R.X, R.X, xx, xnorm));
    ynorm *= lambda;
    return 0;
}

static int vi_solve(p4g_solver_s* s, const double* xl, const double* xu, double* x, double* Fout,
                    p4g_VIReport* rep) {
    const double t0 = now_ms();
    p4g_da_s* da = s->da;
    if (!da->fn) return set_error(PETSC_ERR_ARG_WRONG, "no residual callback: call p4g_da_set_function_local");
    if (!s->snes_type.empty() && s->snes_type != "vinewtonrsls" && s->snes_type != "vinewtonssls")
        return set_error(PETSC_ERR_SUP, "-snes_type %s: p4g_vi_solve implements vinewtonrsls and vinewtonssls",
                         s->snes_type.c_str());
    const bool ssls = s->snes_type == "vinewtonssls";
    if (ssls && (s->pctype == PCT_MG || s->pctype == PCT_JACOBI || s->pctype == PCT_ASM || xu))
        return set_error(PETSC_ERR_SUP, "p4g_vi_solve: vinewtonssls is implemented for lower bounds with -pc_type "
                                        "ilu (the default) or none");
    if (s->ls_type != "bt" && s->ls_type != "basic")
        return set_error(PETSC_ERR_SUP, "-snes_linesearch_type %s is not implemented (bt, basic)", s->ls_type.c_str());
    if (s->mixed) return set_error(PETSC_ERR_SUP, "p4g_vi_solve: mixed precision is not implemented");
    if (s->ksp_rich) return set_error(PETSC_ERR_SUP, "p4g_vi_solve: -ksp_type cg (the default) or gmres");
    if (!da->kcell.empty()) return set_error(PETSC_ERR_SUP, "p4g_vi_solve: variable coefficients are not implemented");
    if (s->fmg) return set_error(PETSC_ERR_SUP, "p4g_vi_solve: -pc_mg_type full is not implemented");
    if (s->coarse_its > 0)
        return set_error(PETSC_ERR_SUP, "p4g_vi_solve: the coarsest reduced system is solved directly (no "
                                        "-mg_coarse_ksp_type cg)");
    if (s->lilu)
        return set_error(PETSC_ERR_SUP, "p4g_vi_solve: ILU smoothing of the Galerkin levels is not implemented");
    if (!s->vi_mode) {
        if (s->ready)
            return set_error(PETSC_ERR_ORDER, "p4g_vi_solve: the solver was set up for linear solves; create a "
                                              "new solver for VI solves");
        s->vi_mode = true;
        if (s->pctype == PCT_MG) s->galerkin = true;   // PCMGSetGalerkin under RSLS
    }
    const double ts = now_ms();
    const bool fresh = !s->ready;
    P4G_TRY(solver_setup(s));
    const double setup_ms = fresh ? now_ms() - ts : 0.0;
    const int L = s->nlev - 1;
    const LevelInfo& FL = s->lv[L];
    const long n = FL.n;
    ViRun R;
    R.s = s;
    R.n = n;
    R.nb = vblocks(n);
    R.st = s->cx.stream;
    R.S = s->cx.S + 16;
    R.g.dim = FL.dim; R.g.mx = FL.mx; R.g.my = FL.my; R.g.mz = FL.mz; R.g.n = n;
    R.g.d = FL.d; R.g.db = FL.db; R.g.sx = FL.sx; R.g.sy = FL.sy; R.g.sz = FL.sz;
    if (!s->vi_buf) {
        P4G_TRY(dalloc(s, (void**)&s->vi_buf, 11 * n * sizeof(double)));
        P4G_TRY(dalloc(s, (void**)&s->vi_act, n));
        P4G_TRY(dalloc(s, (void**)&s->vi_g27, 27 * n * sizeof(double)));
        // the Chebyshev bounds are re-estimated every Newton step: keep their
        // vectors (6 of the finest size) instead of allocating per level
        if (s->sor && s->pctype == PCT_MG && s->lv[L].dim == 2) {   // pipelined strip sweeps on every level
            s->ms_pipes.resize(s->nlev);
            for (int l = 0; l <= L; ++l) {
                if (s->lv[l].mx >= (1 << 22)) continue;
                const int ns = (s->lv[l].my + 63) / 64;
                SorPipe& P = s->ms_pipes[l];
                P4G_TRY(dalloc(s, (void**)&P.flags, (ns + 1) * sizeof(unsigned long long) + sizeof(int)));
                HIP_TRY(hipMemset(P.flags, 0, (ns + 1) * sizeof(unsigned long long) + sizeof(int)));
                P.err = (int*)(P.flags + ns + 1);
                P4G_TRY(dalloc(s, (void**)&P.pk, 10 * s->lv[l].n * sizeof(double)));
                s->lv[l].pipe = &P;
            }
        }
        if (s->pctype == PCT_MG && s->smoother == SM_CHEB) {
            s->est_cap = 6 * n;
            P4G_TRY(dalloc(s, (void**)&s->est_scratch, s->est_cap * sizeof(double)));
        }
    }
    double* b = s->vi_buf;
    R.F0 = b; R.X = b + n; R.F = b + 2 * n; R.Y = b + 3 * n; R.W = b + 4 * n; R.G = b + 5 * n;
    R.Fm = b + 6 * n; R.xl = b + 7 * n; R.xu = xu ? b + 8 * n : nullptr;
    R.ssls = ssls;
    R.phi = b + 9 * n; R.FW = b + 10 * n;
    {   // F(0) by the residual callback: F(x) = J x + F(0) for this linear residual
        std::vector<double> z(n, 0.0), F0(n, 0.0);
        std::vector<double*> r1, r2;
        std::vector<double**> p1, p2;
        void* au = make_view(da, z.data(), r1, p1);
        void* aF = make_view(da, F0.data(), r2, p2);
        DMDALocalInfo info;
        info_of(da, &info);
        P4G_TRY(da->fn(&info, au, aF, da->fnctx));
        HIP_TRY(hipMemcpy(R.F0, F0.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(R.xl, xl, n * sizeof(double), hipMemcpyHostToDevice));
    if (xu) HIP_TRY(hipMemcpy(R.xu, xu, n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(R.W, x, n * sizeof(double), hipMemcpyHostToDevice));
    s->vi_fhist.clear();
    s->vi_kits.clear();
    s->vi_kreason.clear();
    double fnorm, xnorm, ynorm = 0.0, ttol = 0.0, rnorm0 = 0.0;
    if (ssls) {   // F(X), phi(X), |phi| (no projection of the initial iterate)
        HIP_TRY(hipMemcpyAsync(R.X, R.W, n * sizeof(double), hipMemcpyDeviceToDevice, R.st));
        P4G_TRY(R.fun(R.X, R.F, true));
        k_ssls_phi<<<R.nb, 256, 0, R.st>>>(n, R.X, R.xl, R.F, R.phi);
        LAUNCH_CHECK();
        double t;
        P4G_TRY(R.dot(R.phi, R.phi, t, fnorm));
        P4G_TRY(R.dot(R.X, R.X, t, xnorm));
    } else {   // SNESVIProjectOntoBounds, F(X), the inactive-set norm
        k_vi_trial<<<R.nb, 256, 0, R.st>>>(n, R.W, nullptr, 0.0, R.xl, R.xu, R.X);
        LAUNCH_CHECK();
        P4G_TRY(R.fun(R.X, R.F, true));
        P4G_TRY(R.norms(R.X, R.F, fnorm, xnorm));
    }
    p4g_VIReport out;
    std::memset(&out, 0, sizeof out);
    out.fnorm0 = fnorm;
    s->vi_fhist.push_back(fnorm);
    P4G_TRY(sor_pipes_reset(s));
    // PCASM: the engine's preconditioner while this solve runs
    struct AsmScope {
        Engine<double>* e;
        ~AsmScope() { e->asm_op = nullptr; }
    } asm_scope{s->e64.get()};
    if (s->pctype == PCT_ASM) {
        s->e64->asm_op = [s, &R](const double* r, double* z) -> int {
            k_fill<double><<<R.nb, 256, 0, R.st>>>(R.n, 0.0, z);
            if (s->asm_nrows > 0)
                k_asm_apply<<<(int)std::min<long>((s->asm_nrows + 255) / 256, VBLOCKS), 256, 0, R.st>>>(
                    s->asm_nrows, s->asm_row, s->asm_idx, s->asm_inv, r, z);
            LAUNCH_CHECK();
            return 0;
        };
    }
    int reason = vi_converged(s, 0, 0.0, 0.0, fnorm, ttol, rnorm0);
    int i = 0, iter = 0;
    for (; !reason && i < s->snes_maxit; ++i) {
        const double* rhs = R.Fm;
        if (ssls) {   // the semismooth Jacobian at X; Jss Y = phi
            P4G_TRY(ssls_operator(s, R));
            rhs = R.phi;
        } else {      // active set, reduced right-hand side, reduced operator
            k_vi_sets<<<R.nb, 256, 0, R.st>>>(n, R.X, R.F, R.xl, R.xu, s->vi_act, R.Fm, s->cx.part);
            LAUNCH_CHECK();
            P4G_TRY(vi_operator(s, R));
        }
        typename Engine<double>::CGOut co;
        if (s->ksp_gmres)
            P4G_TRY(s->e64->gmres(rhs, R.Y, s->rtol, s->atol, s->dtol, s->maxit, s->gmres_restart, co, s->cx.S));
        else
            P4G_TRY(s->e64->cg(rhs, R.Y, s->normtype, s->rtol, s->atol, s->dtol, s->maxit, co, s->cx.S));
        HIP_TRY(hipStreamSynchronize(R.st));
        P4G_TRY(sor_pipes_check(s, "p4g_vi_solve"));
        s->vi_kits.push_back(co.its);
        s->vi_kreason.push_back(co.reason);
        s->hist = co.hist;
        out.last_ksp_its = co.its;
        out.ksp_its_total += co.its;
        if (co.reason < 0) { reason = -3; break; }   // DIVERGED_LINEAR_SOLVE (maxLinearSolveFailures 1)
        int fail = 0;
        P4G_TRY(vi_line_search(s, R, fnorm, xnorm, ynorm, &fail));
        if (fail) {                                  // maxFailures 1
            reason = (s->snes_stol * xnorm > ynorm) ? 4 : -6;
            break;
        }
        iter = i + 1;
        s->vi_fhist.push_back(fnorm);
        reason = vi_converged(s, iter, xnorm, ynorm, fnorm, ttol, rnorm0);
    }
    if (!reason && i >= s->snes_maxit) reason = -5;   // DIVERGED_MAX_IT
    // the size of the last active set (SSLS: {x <= xl + 1e-8 and F > 0} at the solution)
    if (ssls) {
        std::vector<double> hx(n), hf(n), hl(n);
        HIP_TRY(hipStreamSynchronize(R.st));
        HIP_TRY(hipMemcpy(hx.data(), R.X, n * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(hf.data(), R.F, n * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(hl.data(), R.xl, n * sizeof(double), hipMemcpyDeviceToHost));
        for (long q = 0; q < n; ++q) out.active += (hx[q] <= hl[q] + 1.0e-8 && hf[q] > 0.0) ? 1 : 0;
    } else {
        std::vector<unsigned char> act(n);
        HIP_TRY(hipStreamSynchronize(R.st));
        if (!s->vi_kits.empty()) {
            HIP_TRY(hipMemcpy(act.data(), s->vi_act, n, hipMemcpyDeviceToHost));
            for (long q = 0; q < n; ++q) out.active += act[q];
        }
    }
    HIP_TRY(hipMemcpy(x, R.X, n * sizeof(double), hipMemcpyDeviceToHost));
    if (Fout) HIP_TRY(hipMemcpy(Fout, R.F, n * sizeof(double), hipMemcpyDeviceToHost));
    s->cx.tm.collect();
    out.its = iter;
    out.reason = reason;
    out.fnorm = fnorm;
    out.setup_ms = setup_ms;
    out.solve_ms = now_ms() - t0 - setup_ms;
    if (rep) *rep = out;
    return 0;
}

}  // namespace p4g

extern "C" {

int p4g_vi_solve(p4g_Solver s, const double* xl, const double* xu, double* x, double* F, p4g_VIReport* rep) {
    clear_error();
    if (!s || !xl || !x) return set_error(PETSC_ERR_ARG_WRONG, "p4g_vi_solve: null argument");
    return p4g::vi_solve(s, xl, xu, x, F, rep);
}

int p4g_vi_get_history(p4g_Solver s, double* fnorm, int* ksp_its, int* ksp_reason, int cap, int* nf, int* nk) {
    clear_error();
    if (!s || !nf || !nk) return set_error(PETSC_ERR_ARG_WRONG, "p4g_vi_get_history: null argument");
    *nf = (int)s->vi_fhist.size();
    *nk = (int)s->vi_kits.size();
    for (int q = 0; q < cap && q < *nf; ++q)
        if (fnorm) fnorm[q] = s->vi_fhist[q];
    for (int q = 0; q < cap && q < *nk; ++q) {
        if (ksp_its) ksp_its[q] = s->vi_kits[q];
        if (ksp_reason) ksp_reason[q] = s->vi_kreason[q];
    }
    return 0;
}

}  // extern "C"
