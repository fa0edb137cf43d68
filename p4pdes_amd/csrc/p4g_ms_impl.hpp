// Newton's method for the minimal-surface equation of c/ch7/minimal.c on the
// device: PETSc's SNESSolve_NEWTONLS with SNESLineSearchApply_BT (alpha
// 1e-4) or _Basic and SNESConvergedDefault, the Jacobian by colored finite
// differences (-snes_fd_color, k_ms_fdjac), KSP GMRES (PETSc's default) or
// CG, PC ILU(0) (the serial default), multigrid or none.  Under -pc_type mg
// every coarser level takes the finite-difference Jacobian of the
// rediscretized residual at the injected iterate, as PETSc's PCMG does for a
// DMDA SNES (KSPComputeOperators_SNES with DMRestrictHook_SNESVecSol).  The
// scalar decisions run on the host; every vector pass runs on the device.
// Included once, at the end of p4g_solver.hip.

namespace p4g {

// The coarsest level of a nonsymmetric 27-plane operator: band LU without
// pivoting on the host (PETSc's serial coarse LU factors these diagonally
// dominant Jacobians without row exchanges), then the engine's dense
// inverse on the device, a column per thread (k_band_inverse).
static int coarse_factor_lu(p4g_solver_s* s, const LevelInfo& L, const double* g27, Engine<double>& e,
                            hipStream_t st) {
    const int n = (int)L.n;
    if (n > MAX_COARSE)
        return set_error(PETSC_ERR_SUP, "coarsest grid has %d points (> %d); use more multigrid levels", n,
                         MAX_COARSE);
    if (!e.Ainv || e.n0 != n) return set_error(PETSC_ERR_LIB, "internal: coarse inverse of a different size");
    const int mx = L.mx, my = L.my;
    const int bw = L.dim == 3 ? mx * my + mx + 1 : (L.dim == 2 ? mx + 1 : 1);
    const int w = 2 * bw + 1;
    std::vector<double> g((size_t)27 * n);
    HIP_TRY(hipMemcpy(g.data(), g27, g.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> B((size_t)n * w, 0.0);
    auto at = [&](long r, long c) -> double& { return B[(size_t)r * w + (c - r + bw)]; };
    for (int k = 0; k < L.mz; ++k)
        for (int j = 0; j < my; ++j)
            for (int i = 0; i < mx; ++i) {
                const long q = ((long)k * my + j) * mx + i;
                for (int c = 0; c < 27; ++c) {
                    const int oz = c / 9 - 1, oy = (c / 3) % 3 - 1, ox = c % 3 - 1;
                    if (i + ox < 0 || i + ox >= mx || j + oy < 0 || j + oy >= my || k + oz < 0 || k + oz >= L.mz)
                        continue;
                    at(q, q + ((long)oz * my + oy) * mx + ox) = g[(size_t)c * n + q];
                }
            }
    for (long kk = 0; kk < n; ++kk) {
        const double piv = at(kk, kk);
        if (piv == 0.0) return set_error(PETSC_ERR_LIB, "coarse LU: zero pivot in row %ld", kk);
        const long iend = std::min<long>(kk + bw, n - 1);
        for (long i = kk + 1; i <= iend; ++i) {
            const double l = at(i, kk) / piv;
            if (l == 0.0) continue;
            at(i, kk) = l;
            for (long j = kk + 1; j <= iend; ++j) at(i, j) -= l * at(kk, j);
        }
    }
    if (!s->ms_band || s->ms_band_cap < (long)B.size()) {
        P4G_TRY(dalloc(s, (void**)&s->ms_band, B.size() * sizeof(double)));
        s->ms_band_cap = (long)B.size();
    }
    HIP_TRY(hipMemcpyAsync(s->ms_band, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice, st));
    k_band_inverse<<<(n + 255) / 256, 256, 0, st>>>(n, bw, s->ms_band, e.Ainv);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(st));   // B lives on the host stack
    return 0;
}

struct MsRun {
    p4g_solver_s* s = nullptr;
    long n = 0;
    int nb = 1;
    hipStream_t st = nullptr;
    double *X = nullptr, *F = nullptr, *Y = nullptr, *W = nullptr, *G = nullptr;
    double* S = nullptr;                       // the second half of cx.S
    double *mfw = nullptr, *mfy = nullptr;     // -snes_mf_operator: X + h a, and the product
    std::vector<MsGrid> grid;                  // per level
    std::vector<double*> gb, xl, fl;           // per level: boundary values, injected iterate, its residual

    int sums(double& a, double& b) {
        P4G_TRY(s->cx.finalize(nb, S, FIN_RAW, 0));
        P4G_TRY(s->cx.readS(s->cx.hS));
        a = s->cx.hS[16 + S_TMP0];
        b = s->cx.hS[16 + S_TMP1];
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
    int fun(int l, const double* x, double* f) {
        const MsGrid& g = grid[l];
        k_ms_residual<<<vblocks(g.n), 256, 0, st>>>(g, x, gb[l], f);
        LAUNCH_CHECK();
        return 0;
    }
    // w = x0 - lambda y, gv = F(w), gn = |gv|
    int trial(const double* x0, const double* y, double lambda, double* w, double* gv, double& gn) {
        k_vi_trial<<<nb, 256, 0, st>>>(n, x0, y, lambda, nullptr, nullptr, w);
        LAUNCH_CHECK();
        P4G_TRY(fun((int)grid.size() - 1, w, gv));
        double t;
        return dot(gv, gv, t, gn);
    }
};

// the colored finite-difference Jacobian of level l at x (F = F(x) given)
// into the level's 27-plane table
static int ms_fdjac_level(p4g_solver_s* s, MsRun& R, int l, const double* x, const double* F) {
    const MsGrid& g = R.grid[l];
    const int nb = vblocks(g.n);
    k_vi_dot<<<nb, 256, 0, R.st>>>(g.n, x, x, s->cx.part);
    LAUNCH_CHECK();
    P4G_TRY(s->cx.finalize(nb, R.S, FIN_RAW, 0));
    P4G_TRY(s->cx.readS(s->cx.hS));
    const double unorm = std::sqrt(s->cx.hS[16 + S_TMP1]);
    const double d = 1.0 / (std::sqrt(1.0 + unorm) * 1.490116119384766e-08);   // "wp": PETSC_SQRT_MACHINE_EPSILON
    k_ms_fdjac<<<nb, 256, 0, R.st>>>(g, x, R.gb[l], F, d, 1.0 / d, s->g27[l]);
    if (s->lv[l].pipe) k_sor9_pack<<<nb, 256, 0, R.st>>>(g.n, s->g27[l], s->lv[l].pipe->pk);
    LAUNCH_CHECK();
    return 0;
}

// the registered Jacobian callback's operator on level l (recorded at
// set-up as a constant 5-point stencil: minimal.c registers
// Poisson2DJacobianLocal) into the level's 27-plane table
static int ms_const_level(p4g_solver_s* s, MsRun& R, int l) {
    const LevelInfo& Lv = s->lv[l];
    const int nb = vblocks(Lv.n);
    k_const5_to_g27<<<nb, 256, 0, R.st>>>(Lv.mx, Lv.my, Lv.d, Lv.db, Lv.sx, Lv.sy, s->g27[l]);
    if (Lv.pipe) k_sor9_pack<<<nb, 256, 0, R.st>>>(Lv.n, s->g27[l], Lv.pipe->pk);
    LAUNCH_CHECK();
    return 0;
}

// -snes_mf_operator: y = J(X) a by MatMult_MFFD with the "wp" step
// (MatMFFDCompute_WP, |U| recomputed at every product): h = eps sqrt(1 +
// |X|_2) / |a|_2, y = (F(X + h a) - F(X)) (1/h); a = 0 gives y = 0.  The
// base point is the current iterate and its residual (SNES's MatMFFDSetBase).
static int ms_mffd(p4g_solver_s* s, MsRun& R, double* w, const double* a, double* y) {
    double t, na, nx;
    P4G_TRY(R.dot(a, a, t, na));
    if (na == 0.0) {
        k_fill<double><<<R.nb, 256, 0, R.st>>>(R.n, 0.0, y);
        LAUNCH_CHECK();
        return 0;
    }
    P4G_TRY(R.dot(R.X, R.X, t, nx));
    const double h = 1.490116119384766e-08 * std::sqrt(1.0 + nx) / na;   // PETSC_SQRT_MACHINE_EPSILON
    k_mffd_base<<<R.nb, 256, 0, R.st>>>(R.n, h, a, R.X, w);
    LAUNCH_CHECK();
    P4G_TRY(R.fun((int)R.grid.size() - 1, w, y));
    k_mffd_diff<<<R.nb, 256, 0, R.st>>>(R.n, R.F, 1.0 / h, y);
    LAUNCH_CHECK();
    (void)s;
    return 0;
}

// the Jacobian hierarchy of the current iterate, then the PC's set-up:
// colored finite differences (-snes_fd_color) or the registered Jacobian
static int ms_operator(p4g_solver_s* s, MsRun& R, bool fd) {
    const int L = s->nlev - 1;
    Engine<double>& e = *s->e64;
    if (fd) P4G_TRY(ms_fdjac_level(s, R, L, R.X, R.F));
    else P4G_TRY(ms_const_level(s, R, L));
    if (s->pctype == PCT_ILU) return ilu9_factor(s->lv[L], s->g27[L], s->ms_ilu9, R.st);
    if (s->pctype != PCT_MG) return 0;
    for (int l = L - 1; l >= 0; --l) {
        if (!fd) {
            P4G_TRY(ms_const_level(s, R, l));
            continue;
        }
        const MsGrid &gc = R.grid[l], &gf = R.grid[l + 1];
        const double* xf = (l + 1 == L) ? R.X : R.xl[l + 1];
        k_inject2<<<vblocks(gc.n), 256, 0, R.st>>>(gc.mx, gc.my, gf.mx, xf, R.xl[l]);
        LAUNCH_CHECK();
        P4G_TRY(R.fun(l, R.xl[l], R.fl[l]));
        P4G_TRY(ms_fdjac_level(s, R, l, R.xl[l], R.fl[l]));
    }
    if (L > 0) {
        HIP_TRY(hipStreamSynchronize(R.st));
        P4G_TRY(coarse_factor_lu(s, s->lv[0], s->g27[0], e, R.st));
    }
    if (s->smoother == SM_CHEB)   // one level: the level smoother is the preconditioner
        for (int l = L > 0 ? 1 : 0; l <= L; ++l) P4G_TRY(estimate_level(s, l));
    return 0;
}

// MatIsSymmetric(J, tol) of the finest Jacobian (-mat_is_symmetric)
static int ms_symmetric(p4g_solver_s* s, MsRun& R, double tol, int* sym) {
    const LevelInfo& F = s->lv.back();
    const int nb = vblocks(F.n);
    k_sym27<<<nb, 256, 0, R.st>>>(F.mx, F.my, s->g27.back(), s->cx.part);
    LAUNCH_CHECK();
    std::vector<double> part(nb);
    HIP_TRY(hipMemcpyAsync(part.data(), s->cx.part, nb * sizeof(double), hipMemcpyDeviceToHost, R.st));
    HIP_TRY(hipStreamSynchronize(R.st));
    double m = 0.0;
    for (double v : part) m = std::max(m, v);
    *sym = m <= tol ? 1 : 0;
    return 0;
}

// SNESLineSearchApply_BT with alpha = 1e-4 (or _Basic).  On return X, F hold
// the accepted point (pointers swapped with the trial buffers), fnorm its
// norm, ynorm = lambda |Y|; *fail = 1 when the search failed.
static int ms_line_search(p4g_solver_s* s, MsRun& R, double& fnorm, double& xnorm, double& ynorm, int* fail) {
    const double alpha = 1e-4, steptol = 1e-12;
    double lambda = 1.0, t;
    *fail = 0;
    P4G_TRY(R.dot(R.X, R.X, t, xnorm));
    P4G_TRY(R.dot(R.Y, R.Y, t, ynorm));
    double gnorm = 0.0;
    if (s->ls_type == "basic") {
        P4G_TRY(R.trial(R.X, R.Y, 1.0, R.W, R.G, gnorm));
    } else {
        if (ynorm == 0.0) { *fail = 1; return 0; }
        P4G_TRY(s->e64->mat_mult(R.Y, R.W));   // initial slope of |F|^2/2 along -Y: -(F, J Y)
        double initslope, wn;
        P4G_TRY(R.dot(R.F, R.W, initslope, wn));
        if (initslope > 0.0) initslope = -initslope;
        if (initslope == 0.0) initslope = -1.0;
        const double f = fnorm * fnorm;
        P4G_TRY(R.trial(R.X, R.Y, lambda, R.W, R.G, gnorm));
        double g = gnorm * gnorm;
        bool ok = 0.5 * g <= 0.5 * f + lambda * alpha * initslope;
        if (!ok) {
            std::vector<double> hx(R.n), hy(R.n);
            HIP_TRY(hipStreamSynchronize(R.st));
            HIP_TRY(hipMemcpy(hx.data(), R.X, R.n * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hy.data(), R.Y, R.n * sizeof(double), hipMemcpyDeviceToHost));
            double rellength = 0.0;
            for (long q = 0; q < R.n; ++q)
                rellength = std::max(rellength, std::fabs(hy[q]) / std::max(std::fabs(hx[q]), 1.0));
            const double minlambda = steptol / rellength;
            double lt = -initslope / (g - f - 2.0 * lambda * initslope);   // quadratic model
            double lprev = lambda, gprev = g;
            if (lt > 0.5 * lambda) lt = 0.5 * lambda;
            lambda = lt <= 0.1 * lambda ? 0.1 * lambda : lt;
            P4G_TRY(R.trial(R.X, R.Y, lambda, R.W, R.G, gnorm));
            g = gnorm * gnorm;
            ok = 0.5 * g < 0.5 * f + lambda * alpha * initslope;
            while (!ok) {   // cubic model
                if (lambda <= minlambda) break;
                const double t1 = 0.5 * (g - f) - lambda * initslope;
                const double t2 = 0.5 * (gprev - f) - lprev * initslope;
                const double a = (t1 / (lambda * lambda) - t2 / (lprev * lprev)) / (lambda - lprev);
                const double b = (-lprev * t1 / (lambda * lambda) + lambda * t2 / (lprev * lprev)) / (lambda - lprev);
                double dd = b * b - 3.0 * a * initslope;
                if (dd < 0.0) dd = 0.0;
                lt = (a == 0.0) ? -initslope / (2.0 * b) : (-b + std::sqrt(dd)) / (3.0 * a);
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
    std::swap(R.F, R.G);
    fnorm = gnorm;
    P4G_TRY(R.dot(R.X, R.X, t, xnorm));
    ynorm *= lambda;
    return 0;
}

static int ms_solve(p4g_solver_s* s, double q, const double* gb, double* x, p4g_MSMonitorFn mon, void* monctx,
                    p4g_MSReport* rep) {
    const double t0 = now_ms();
    p4g_da_s* da = s->da;
    if (da->dim != 2) return set_error(PETSC_ERR_SUP, "p4g_ms_solve: the minimal-surface equation is 2D");
    if (!s->snes_type.empty() && s->snes_type != "newtonls")
        return set_error(PETSC_ERR_SUP, "-snes_type %s: p4g_ms_solve implements newtonls", s->snes_type.c_str());
    if (s->ls_type != "bt" && s->ls_type != "basic")
        return set_error(PETSC_ERR_SUP, "-snes_linesearch_type %s is not implemented (bt, basic)", s->ls_type.c_str());
    if (s->opt.has("snes_mf"))
        return set_error(PETSC_ERR_SUP, "p4g_ms_solve: -snes_mf (no preconditioning matrix) is not implemented; "
                                        "use -snes_fd_color or -snes_mf_operator");
    // the preconditioning matrix: colored finite differences, else the
    // registered Jacobian; the operator: matrix-free under -snes_mf_operator
    const bool fd = s->opt.has("snes_fd_color"), mf = s->opt.has("snes_mf_operator");
    if (s->mixed) return set_error(PETSC_ERR_SUP, "p4g_ms_solve: mixed precision is not implemented");
    if (s->fmg || s->kaskade || s->mcycles != 1 || s->coarse_its > 0 || s->lilu || s->galerkin)
        return set_error(PETSC_ERR_SUP, "p4g_ms_solve: multiplicative V/W cycles over the finite-difference "
                                        "Jacobians, with a direct coarse solve and Chebyshev/Richardson smoothing "
                                        "(Jacobi or SOR)");
    if (s->ksp_rich) return set_error(PETSC_ERR_SUP, "p4g_ms_solve: -ksp_type gmres (the default) or cg");
    if (s->pctype == PCT_JACOBI || s->pctype == PCT_ASM)
        return set_error(PETSC_ERR_SUP, "p4g_ms_solve: -pc_type ilu (the default), mg or none");
    const bool gmres = s->ksp_gmres || !s->opt.has("ksp_type");   // SNES's KSP defaults to GMRES
    if (!s->ms_mode) {
        if (s->ready)
            return set_error(PETSC_ERR_ORDER, "p4g_ms_solve: the solver was set up for other solves; create a new "
                                              "solver for minimal-surface solves");
        s->ms_mode = true;
        s->vi_mode = true;   // operators formed per Newton step, not at set-up
    }
    const double ts = now_ms();
    const bool fresh = !s->ready;
    P4G_TRY(solver_setup(s));
    const int L = s->nlev - 1;
    const long n = s->lv[L].n;
    const int l0 = s->pctype == PCT_MG ? 0 : L;   // levels that take an operator
    for (int l = l0; l < L; ++l)
        if (s->lv[l + 1].mx != 2 * s->lv[l].mx - 1 || s->lv[l + 1].my != 2 * s->lv[l].my - 1)
            return set_error(PETSC_ERR_SUP, "p4g_ms_solve: multigrid needs vertex-nested levels (2 m - 1)");
    if ((int)s->g27.size() != s->nlev) s->g27.assign(s->nlev, nullptr);
    if ((int)s->e64->g27.size() != s->nlev) s->e64->g27.assign(s->nlev, nullptr);
    MsRun R;
    R.s = s;
    R.n = n;
    R.nb = vblocks(n);
    R.st = s->cx.stream;
    R.S = s->cx.S + 16;
    R.grid.resize(s->nlev);
    R.gb.assign(s->nlev, nullptr);
    R.xl.assign(s->nlev, nullptr);
    R.fl.assign(s->nlev, nullptr);
    for (int l = 0; l <= L; ++l) {
        const LevelInfo& Lv = s->lv[l];
        R.grid[l] = MsGrid{Lv.mx, Lv.my, Lv.n, 1.0 / (Lv.mx - 1), 1.0 / (Lv.my - 1), q};
    }
    if (!s->ms_buf) {
        long tot = 7 * n;   // X F Y W G, and the matrix-free product's two vectors
        for (int l = l0; l <= L; ++l) tot += (l == L ? 1 : 3) * s->lv[l].n;
        P4G_TRY(dalloc(s, (void**)&s->ms_buf, tot * sizeof(double)));
        for (int l = l0; l <= L; ++l) {
            const size_t bytes = 27 * s->lv[l].n * sizeof(double);
            P4G_TRY(dalloc(s, (void**)&s->g27[l], bytes));
            HIP_TRY(hipMemset(s->g27[l], 0, bytes));   // the 2D operator uses planes 9..17 only
            s->lv[l].gal = true;
            s->e64->g27[l] = s->g27[l];
        }
        if (s->sor && s->pctype == PCT_MG) {   // pipelined strip sweeps on every level
            s->ms_pipes.resize(s->nlev);
            for (int l = l0; l <= L; ++l) {
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
        if (s->pctype == PCT_ILU) {
            P4G_TRY(dalloc(s, (void**)&s->ms_ilu9, 9 * n * sizeof(double)));
            s->e64->ilu9 = s->ms_ilu9;
        }
        if (s->pctype == PCT_MG && s->smoother == SM_CHEB) {
            s->est_cap = 6 * n;
            P4G_TRY(dalloc(s, (void**)&s->est_scratch, s->est_cap * sizeof(double)));
        }
    }
    {
        double* b = s->ms_buf;
        R.X = b; R.F = b + n; R.Y = b + 2 * n; R.W = b + 3 * n; R.G = b + 4 * n;
        R.mfw = b + 5 * n; R.mfy = b + 6 * n;
        b += 7 * n;
        for (int l = L; l >= l0; --l) {
            R.gb[l] = b;
            b += s->lv[l].n;
            if (l < L) { R.xl[l] = b; R.fl[l] = b + s->lv[l].n; b += 2 * s->lv[l].n; }
        }
    }
    HIP_TRY(hipMemcpy(R.gb[L], gb, n * sizeof(double), hipMemcpyHostToDevice));
    for (int l = L - 1; l >= l0; --l) {   // coarse boundary values: the coincident fine nodes
        k_inject2<<<vblocks(s->lv[l].n), 256, 0, R.st>>>(s->lv[l].mx, s->lv[l].my, s->lv[l + 1].mx, R.gb[l + 1],
                                                         R.gb[l]);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipMemcpy(R.X, x, n * sizeof(double), hipMemcpyHostToDevice));
    const double setup_ms = fresh ? now_ms() - ts : 0.0;
    P4G_TRY(R.fun(L, R.X, R.F));
    if (da->fn) {   // the residual callback must be the operator the device evaluates
        std::vector<double> Fh(n, 0.0), Fd(n), xh(x, x + n);
        std::vector<double*> r1, r2;
        std::vector<double**> p1, p2;
        void* au = make_view(da, xh.data(), r1, p1);
        void* aF = make_view(da, Fh.data(), r2, p2);
        DMDALocalInfo info;
        info_of(da, &info);
        P4G_TRY(da->fn(&info, au, aF, da->fnctx));
        HIP_TRY(hipMemcpy(Fd.data(), R.F, n * sizeof(double), hipMemcpyDeviceToHost));
        double dmax = 0.0, fmax = 0.0;
        for (long k = 0; k < n; ++k) {
            dmax = std::max(dmax, std::fabs(Fd[k] - Fh[k]));
            fmax = std::max(fmax, std::fabs(Fh[k]));
        }
        if (!(dmax <= 1e-10 * std::max(1.0, fmax)))
            return set_error(PETSC_ERR_ARG_WRONG, "p4g_ms_solve: the residual callback differs from the minimal-surface "
                                                  "residual with q = %g and these boundary values (by %.3e)", q, dmax);
    }
    s->ms_fhist.clear();
    s->ms_kits.clear();
    s->ms_kreason.clear();
    s->ms_sym.clear();
    const bool symcheck = s->opt.has("mat_is_symmetric");
    const double symtol = s->opt.getr("mat_is_symmetric", 0.0);
    std::vector<double> xh;
    auto monitor = [&](int it, double fn) -> int {
        if (!mon) return 0;
        xh.resize(n);
        HIP_TRY(hipMemcpy(xh.data(), R.X, n * sizeof(double), hipMemcpyDeviceToHost));
        const int e = mon(monctx, it, fn, xh.data());
        if (e) return set_error(e, "p4g_ms_solve: the monitor returned %d", e);
        return 0;
    };
    double fnorm, xnorm, ynorm = 0.0, ttol = 0.0, rnorm0 = 0.0, t;
    P4G_TRY(R.dot(R.F, R.F, t, fnorm));
    p4g_MSReport out;
    std::memset(&out, 0, sizeof out);
    out.fnorm0 = fnorm;
    s->ms_fhist.push_back(fnorm);
    P4G_TRY(monitor(0, fnorm));
    P4G_TRY(sor_pipes_reset(s));
    int reason = vi_converged(s, 0, 0.0, 0.0, fnorm, ttol, rnorm0);
    int i = 0, iter = 0;
    // the engine's finest operator is the matrix-free product while this
    // solve runs (R holds its base point)
    struct MfScope {
        Engine<double>* e;
        ~MfScope() { e->mf_op = nullptr; e->mf_tmp = nullptr; }
    } mf_scope{s->e64.get()};
    if (mf) {
        s->e64->mf_tmp = R.mfy;
        s->e64->mf_op = [s, &R](const double* a, double* y) { return ms_mffd(s, R, R.mfw, a, y); };
    }
    for (; !reason && i < s->snes_maxit; ++i) {
        P4G_TRY(ms_operator(s, R, fd));
        if (symcheck) {
            int sym = 0;
            P4G_TRY(ms_symmetric(s, R, symtol, &sym));
            s->ms_sym.push_back(sym);
        }
        typename Engine<double>::CGOut co;
        if (gmres)
            P4G_TRY(s->e64->gmres(R.F, R.Y, s->rtol, s->atol, s->dtol, s->maxit, s->gmres_restart, co, s->cx.S));
        else
            P4G_TRY(s->e64->cg(R.F, R.Y, s->normtype, s->rtol, s->atol, s->dtol, s->maxit, co, s->cx.S));
        HIP_TRY(hipStreamSynchronize(R.st));
        P4G_TRY(sor_pipes_check(s, "p4g_ms_solve"));
        s->ms_kits.push_back(co.its);
        s->ms_kreason.push_back(co.reason);
        s->hist = co.hist;
        out.last_ksp_its = co.its;
        out.ksp_its_total += co.its;
        if (co.reason < 0) { reason = -3; break; }   // DIVERGED_LINEAR_SOLVE (maxLinearSolveFailures 1)
        int fail = 0;
        P4G_TRY(ms_line_search(s, R, fnorm, xnorm, ynorm, &fail));
        if (fail) {                                  // maxFailures 1
            reason = (s->snes_stol * xnorm > ynorm) ? 4 : -6;
            break;
        }
        iter = i + 1;
        s->ms_fhist.push_back(fnorm);
        P4G_TRY(monitor(iter, fnorm));
        reason = vi_converged(s, iter, xnorm, ynorm, fnorm, ttol, rnorm0);
    }
    if (!reason && i >= s->snes_maxit) reason = -5;   // DIVERGED_MAX_IT
    HIP_TRY(hipMemcpy(x, R.X, n * sizeof(double), hipMemcpyDeviceToHost));
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

int p4g_ms_solve(p4g_Solver s, double q, const double* gb, double* x, p4g_MSMonitorFn mon, void* monctx,
                 p4g_MSReport* rep) {
    clear_error();
    if (!s || !gb || !x) return set_error(PETSC_ERR_ARG_WRONG, "p4g_ms_solve: null argument");
    return p4g::ms_solve(s, q, gb, x, mon, monctx, rep);
}

int p4g_ms_get_history(p4g_Solver s, double* fnorm, int* ksp_its, int* ksp_reason, int* jac_sym, int cap, int* nf,
                       int* nk) {
    clear_error();
    if (!s || !nf || !nk) return set_error(PETSC_ERR_ARG_WRONG, "p4g_ms_get_history: null argument");
    *nf = (int)s->ms_fhist.size();
    *nk = (int)s->ms_kits.size();
    for (int k = 0; k < cap && k < *nf; ++k)
        if (fnorm) fnorm[k] = s->ms_fhist[k];
    for (int k = 0; k < cap && k < *nk; ++k) {
        if (ksp_its) ksp_its[k] = s->ms_kits[k];
        if (ksp_reason) ksp_reason[k] = s->ms_kreason[k];
        if (jac_sym) jac_sym[k] = k < (int)s->ms_sym.size() ? s->ms_sym[k] : -1;
    }
    return 0;
}

}  // extern "C"
