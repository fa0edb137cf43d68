/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * c/ch7/minimal.c restated: the minimal surface equation
 *   -div((1 + |grad u|^2)^q grad u) = 0   on (0,1)^2,  u = g on the boundary,
 * on a box-stencil DMDA (minimal.c:141-148), discretized by minimal.c's
 * FormFunctionLocal (minimal.c:230-301), solved by Newton's method with
 * the Jacobian by finite differences with coloring (-snes_fd_color), with
 * optional grid sequencing and the area monitor (minimal.c:305-362).
 *
 * PETSc semantics restated (PETSc is not vendored; no version is pinned):
 *   - SNESSolve_NEWTONLS: F(x), |F|, per iteration J(x), KSPSolve(J, F) for
 *     the direction Y (zero initial guess), the line search x <- x - lambda Y,
 *     SNESConvergedDefault; one linear-solve or line-search failure ends the
 *     solve (maxLinearSolveFailures = maxFailures = 1);
 *   - SNESLineSearchApply_BT with alpha = 1e-4: full step, then quadratic and
 *     cubic backtracking (Dennis & Schnabel, A6.3.1);
 *   - MatFDColoringApply with the default "wp" step: h = 1/d with
 *     d = 1/(sqrt(1 + |x|_2) * 1.490116119384766e-08), w3 = x + h on the
 *     columns of one color, J(row, col) = (F(w3) - F(x))_row * d; the DMDA
 *     box-stencil coloring (i mod 3) + 3 (j mod 3) and the DMDA box pattern
 *     (explicit zeros kept, which shapes ILU(0));
 *   - KSPGMRES (oracle_solver.c) and PCILU(0) are the serial defaults;
 *   - PCMG from the DMDA hierarchy: on each coarser level the solution is
 *     injected and the level operator is the colored finite-difference
 *     Jacobian of the rediscretized residual there (KSPComputeOperators_SNES
 *     with DMRestrictHook_SNESVecSol); Pᵀ restriction, Q1 interpolation.
 *     Chebyshev smoothing estimates its interval with the CG-Lanczos
 *     estimate of oracle_solver.c; PETSc's estimator for a nonsymmetric
 *     operator may differ: parity with PETSc unpinned for -pc_type mg.
 *
 *   - -snes_mf_operator: the Krylov operator is MatMFFD ("wp" step) at the
 *     current iterate; without -snes_fd_color the preconditioning matrix on
 *     every level is the registered Poisson2DJacobianLocal (minimal.c:159-
 *     160); PCMG's finest level smooths with the matrix-free operator and
 *     SOR on the Poisson matrix; -oracle_nranks N simulates N processes'
 *     y-split DMDA (local SOR sweeps on each level).
 *
 * Pinning: minimal.test1 (the |F| history 1.08276 .. 1.772e-10 of Newton /
 * GMRES / ILU on the 5x5 catenoid, the reason and the error line),
 * minimal.test2 (q = 0: "Matrix is symmetric" lines and CG/ILU iteration
 * counts 5 and 6 on the 9x9 tent), minimal.test3 (-snes_mf_operator with
 * PCMG on two processes, grid-sequenced 3x3 -> 9x9: every area/D line of
 * -ms_monitor and the Newton counts 5, 3, 3) and minimal.test4 (grid-
 * sequenced tent: Newton counts 3, 5, 5).  tests/test_minimal.py also checks the
 * finite-difference Jacobian against the analytic derivative of the
 * residual and, for q = 0, against the 5-point Laplacian.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"
#include "oracle_internal.h"

enum { MS_TENT = 0, MS_CATENOID = 1 };

typedef struct {
    int    problem, quaddeg;
    double q, H, c;
} msctx;

/* boundary values: minimal.c:29-45 */
static double ms_g(const msctx *m, double x, double y) {
    if (m->problem == MS_TENT) return x < 1.0e-8 ? 2.0 * m->H * (y < 0.5 ? y : 1.0 - y) : 0.0;
    const double ch = cosh(x / m->c);
    return m->c * ch * sin(acos((y / m->c) / ch));
}

static double ms_D(double w, double q) { return pow(1.0 + w, q); }   /* minimal.c:49-51 */

/* FormFunctionLocal (minimal.c:230-301) at node (i, j) */
static double ms_point(const ogrid *g, const msctx *m, const double *u, int i, int j) {
    const int mx = g->mx, my = g->my;
    const double hx = 1.0 / (mx - 1), hy = 1.0 / (my - 1);
    const double x = i * hx, y = j * hy;
    const long q = (long)j * mx + i;
    if (j == 0 || i == 0 || i == mx - 1 || j == my - 1) return u[q] - ms_g(m, x, y);
    const double uc = u[q];
    const double ue = (i + 1 == mx - 1) ? ms_g(m, x + hx, y) : u[q + 1];
    const double uw = (i - 1 == 0) ? ms_g(m, x - hx, y) : u[q - 1];
    const double un = (j + 1 == my - 1) ? ms_g(m, x, y + hy) : u[q + mx];
    const double us = (j - 1 == 0) ? ms_g(m, x, y - hy) : u[q - mx];
    const double une = (i + 1 == mx - 1 || j + 1 == my - 1) ? ms_g(m, x + hx, y + hy) : u[q + mx + 1];
    const double unw = (i - 1 == 0 || j + 1 == my - 1) ? ms_g(m, x - hx, y + hy) : u[q + mx - 1];
    const double use = (i + 1 == mx - 1 || j - 1 == 0) ? ms_g(m, x + hx, y - hy) : u[q - mx + 1];
    const double usw = (i - 1 == 0 || j - 1 == 0) ? ms_g(m, x - hx, y - hy) : u[q - mx - 1];
    double dux, duy;
    dux = (ue - uc) / hx;
    duy = (un + une - us - use) / (4.0 * hy);
    const double De = ms_D(dux * dux + duy * duy, m->q);
    dux = (uc - uw) / hx;
    duy = (unw + un - usw - us) / (4.0 * hy);
    const double Dw = ms_D(dux * dux + duy * duy, m->q);
    dux = (ue + une - uw - unw) / (4.0 * hx);
    duy = (un - uc) / hy;
    const double Dn = ms_D(dux * dux + duy * duy, m->q);
    dux = (ue + use - uw - usw) / (4.0 * hx);
    duy = (uc - us) / hy;
    const double Ds = ms_D(dux * dux + duy * duy, m->q);
    return -(hy / hx) * (De * (ue - uc) - Dw * (uc - uw)) - (hx / hy) * (Dn * (un - uc) - Ds * (uc - us));
}

static void ms_residual(const ogrid *g, const msctx *m, const double *u, double *F) {
    for (int j = 0; j < g->my; j++)
        for (int i = 0; i < g->mx; i++) F[(long)j * g->mx + i] = ms_point(g, m, u, i, j);
}

/* the DMDA box-stencil (width 1) pattern, columns ascending */
static void box_pattern(const ogrid *g, ocsr *J) {
    const int mx = g->mx, my = g->my;
    const long n = g->n;
    J->n = n;
    J->rp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    J->ci = (int32_t *)malloc(sizeof(int32_t) * 9 * n);
    J->v = (double *)calloc(9 * n, sizeof(double));
    int64_t w = 0;
    J->rp[0] = 0;
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            for (int oy = -1; oy <= 1; oy++)
                for (int ox = -1; ox <= 1; ox++) {
                    const int ii = i + ox, jj = j + oy;
                    if (ii < 0 || ii >= mx || jj < 0 || jj >= my) continue;
                    J->ci[w++] = (int32_t)((long)jj * mx + ii);
                }
            J->rp[(long)j * mx + i + 1] = w;
        }
}

#define FD_EPS 1.490116119384766e-08   /* PETSC_SQRT_MACHINE_EPSILON */

/* MatFDColoringApply_AIJ, "wp" step, DMDA box coloring; F0 = F(x) */
static void ms_fdjac(const ogrid *g, const msctx *m, const double *x, const double *F0, ocsr *J) {
    const int mx = g->mx, my = g->my;
    const long n = g->n;
    box_pattern(g, J);
    const double unorm = ovnorm2(n, x);
    const double d = 1.0 / (sqrt(1.0 + unorm) * FD_EPS);
    const double h = 1.0 / d;
    double *w3 = (double *)malloc(sizeof(double) * n), *w2 = (double *)malloc(sizeof(double) * n);
    for (int color = 0; color < 9; color++) {
        const int ci = color % 3, cj = color / 3;
        memcpy(w3, x, sizeof(double) * n);
        for (int j = cj; j < my; j += 3)
            for (int i = ci; i < mx; i += 3) w3[(long)j * mx + i] += h;
        ms_residual(g, m, w3, w2);
        for (long r = 0; r < n; r++) w2[r] -= F0[r];
        /* each row has at most one column of this color in its box */
        for (long r = 0; r < n; r++)
            for (int64_t p = J->rp[r]; p < J->rp[r + 1]; p++) {
                const long c = J->ci[p];
                if ((int)(c % mx) % 3 == ci && (int)(c / mx) % 3 == cj) J->v[p] = w2[r] * d;
            }
    }
    free(w3); free(w2);
}

/* ------------------------------------------------ -snes_mf_operator */

/* MatMFFD at the base point (U, F(U)) with the default "wp" step
 * (MatMFFDCompute_WP, MatMult_MFFD): h = eps sqrt(1 + |U|_2) / |a|_2,
 * y = (F(U + h a) - F(U)) * (1/h); a = 0 gives y = 0. */
typedef struct {
    const ogrid *g;
    const msctx *m;
    const double *U, *F;
    double *w;
    long nmult;
} msmf;

static void msmf_mult(void *ctx, const double *a, double *y) {
    msmf *c = (msmf *)ctx;
    const long n = c->g->n;
    c->nmult++;
    const double norma = ovnorm2(n, a);
    if (norma == 0.0) { ovset(n, 0.0, y); return; }
    const double h = FD_EPS * sqrt(1.0 + ovnorm2(n, c->U)) / norma;
    for (long i = 0; i < n; i++) c->w[i] = h * a[i] + c->U[i];
    ms_residual(c->g, c->m, c->w, y);
    const double ih = 1.0 / h;
    for (long i = 0; i < n; i++) y[i] = (y[i] - c->F[i]) * ih;
}

/* the Jacobian that minimal.c registers (Poisson2DJacobianLocal,
 * c/ch6/poissonfunctions.c:152-193, cx = cy = 1): every row's diagonal
 * 2 (hy/hx + hx/hy), interior rows coupled to interior neighbours by
 * -hy/hx and -hx/hy; on the box DMDA's 9-point pattern with the other
 * entries explicit zeros, as DMCreateMatrix preallocates them (this shapes
 * ILU(0)) */
static void ms_poisson(const ogrid *g, ocsr *J) {
    const int mx = g->mx, my = g->my;
    const double hx = 1.0 / (mx - 1), hy = 1.0 / (my - 1);
    const double scx = hy / hx, scy = hx / hy, scd = 2.0 * (scx + scy);
    box_pattern(g, J);
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            const long q = (long)j * mx + i;
            const int in = i > 0 && i < mx - 1 && j > 0 && j < my - 1;
            for (int64_t p = J->rp[q]; p < J->rp[q + 1]; p++) {
                const long c = J->ci[p];
                const int ci = (int)(c % mx), cj = (int)(c / mx);
                const int cin = ci > 0 && ci < mx - 1 && cj > 0 && cj < my - 1;
                if (c == q) J->v[p] = scd;
                else if (in && cin && cj == j && (ci == i - 1 || ci == i + 1)) J->v[p] = -scx;
                else if (in && cin && ci == i && (cj == j - 1 || cj == j + 1)) J->v[p] = -scy;
            }
        }
}

/* ------------------------------------------- PCMG over colored Jacobians */

typedef struct {
    int     nlev, cycles;
    ogrid  *g;
    ocsr   *A, *P;        /* A[nlev-1] borrowed; P[l]: level l-1 -> l       */
    oksp   *sm;
    double **x, **b, **r;
    msmf   *mf;           /* -snes_mf_operator: the finest level's operator */
} msmg;

/* r = b - A x on level l (the matrix-free operator on the finest level) */
static void msmg_resid(msmg *m, int l, const double *b, const double *x, double *r) {
    if (m->mf && l == m->nlev - 1) {
        msmf_mult(m->mf, x, r);
        for (long i = 0; i < m->g[l].n; i++) r[i] = b[i] - r[i];
    } else {
        ocsr_resid(&m->A[l], b, x, r);
    }
}

static void msmg_free(msmg *m) {
    for (int l = 0; l < m->nlev; l++) {
        oksp_free(&m->sm[l]);
        if (l < m->nlev - 1) { ocsr_free(&m->A[l]); free(m->x[l]); free(m->b[l]); }
        if (m->P[l].v) ocsr_free(&m->P[l]);
        free(m->r[l]);
    }
    free(m->g); free(m->A); free(m->P); free(m->sm); free(m->x); free(m->b); free(m->r);
    memset(m, 0, sizeof *m);
}

static int msmg_setup(msmg *m, const ogrid *fine, const msctx *mc, int nlev, const double *X, const ocsr *J,
                      int fd, msmf *mf, int nranks, oopts *o, char *err, int errlen) {
    memset(m, 0, sizeof *m);
    m->nlev = nlev;
    m->mf = mf;
    m->g = (ogrid *)calloc(nlev, sizeof(ogrid));
    m->A = (ocsr *)calloc(nlev, sizeof(ocsr));
    m->P = (ocsr *)calloc(nlev, sizeof(ocsr));
    m->sm = (oksp *)calloc(nlev, sizeof(oksp));
    m->x = (double **)calloc(nlev, sizeof(double *));
    m->b = (double **)calloc(nlev, sizeof(double *));
    m->r = (double **)calloc(nlev, sizeof(double *));
    const char *ct = oopts_get(o, "pc_mg_cycle_type");
    m->cycles = (ct && !strcmp(ct, "w")) ? 2 : 1;
    const int L = nlev - 1;
    if (nranks > 1 && nlev > 16) {
        snprintf(err, errlen, "oracle: simulated ranks with more than 16 levels");
        return 63;
    }
    m->g[L] = *fine;
    for (int l = L; l > 0; l--)
        if (ogrid_coarsen(&m->g[l], &m->g[l - 1])) {
            snprintf(err, errlen, "oracle: grid of %d levels cannot be coarsened", nlev);
            return 63;
        }
    m->A[L] = *J;
    double *Xf = (double *)malloc(sizeof(double) * fine->n);
    memcpy(Xf, X, sizeof(double) * fine->n);
    for (int l = L; l > 0; l--) {
        const ogrid *gf = &m->g[l], *gc = &m->g[l - 1];
        oq1_build_P(gc, gf, &m->P[l]);
        /* injection of the solution, then the colored Jacobian there (without
         * -snes_fd_color the registered Poisson Jacobian, which does not
         * depend on the iterate) */
        double *Xc = (double *)malloc(sizeof(double) * gc->n), *Fc = (double *)malloc(sizeof(double) * gc->n);
        for (int j = 0; j < gc->my; j++)
            for (int i = 0; i < gc->mx; i++) Xc[(long)j * gc->mx + i] = Xf[(long)(2 * j) * gf->mx + 2 * i];
        if (fd) {
            ms_residual(gc, mc, Xc, Fc);
            ms_fdjac(gc, mc, Xc, Fc, &m->A[l - 1]);
        } else {
            ms_poisson(gc, &m->A[l - 1]);
        }
        free(Xf); free(Fc);
        Xf = Xc;
    }
    free(Xf);
    int ltype = oracle_parse_ksp(oopts_get(o, "mg_levels_ksp_type"), KSP_CHEBYSHEV);
    int lpc = oracle_parse_pc(oopts_get(o, "mg_levels_pc_type"), PC_SOR);
    int lits = oopts_int(o, "mg_levels_ksp_max_it", 2);
    double lscale = oopts_real(o, "mg_levels_ksp_richardson_scale", 1.0);
    int ctype = oracle_parse_ksp(oopts_get(o, "mg_coarse_ksp_type"), KSP_PREONLY);
    int cpc = oracle_parse_pc(oopts_get(o, "mg_coarse_pc_type"), PC_LU);
    int esteig = oopts_int(o, "mg_levels_ksp_chebyshev_esteig_steps", 10);
    if (ltype < 0 || lpc < 0 || lpc == PC_MG || ltype == KSP_GMRES || ctype < 0 || cpc < 0 || cpc == PC_MG) {
        snprintf(err, errlen, "oracle: unsupported level smoother options");
        return 56;
    }
    /* simulated ranks: the DMDA's y-split of every level (rank q owns rows
     * split[q] .. split[q+1]-1; a coarse row I belongs to the owner of fine
     * row 2I), SOR sweeps local to each rank's rows */
    int split[16][17];
    if (nranks > 16) nranks = 16;
    if (nranks > 1) {
        split[L][0] = 0;
        for (int q = 0; q < nranks; q++)
            split[L][q + 1] = split[L][q] + m->g[L].my / nranks + (m->g[L].my % nranks > q);
        for (int l = L; l > 0; l--) {
            split[l - 1][0] = 0;
            for (int q = 0; q < nranks; q++) {
                int cnt = 0;
                for (int I = 0; I < m->g[l - 1].my; I++)
                    if (2 * I >= split[l][q] && 2 * I < split[l][q + 1]) cnt++;
                split[l - 1][q + 1] = split[l - 1][q] + cnt;
            }
        }
    }
    for (int l = 0; l < nlev; l++) {
        const long n = m->g[l].n;
        if (l > 0 || nlev == 1) {
            oksp_init(&m->sm[l], ltype, &m->A[l], lpc);
            if (mf && l == L) { m->sm[l].amult = msmf_mult; m->sm[l].actx = mf; }
            if (nranks > 1 && lpc == PC_SOR && l < 16) {
                m->sm[l].pc.nblk = nranks;
                for (int q = 0; q <= nranks; q++) m->sm[l].pc.blk[q] = (long)split[l][q] * m->g[l].mx;
            }
            m->sm[l].maxit = lits;
            m->sm[l].rich_scale = lscale;
            m->sm[l].esteig_steps = esteig < 1 ? 1 : (esteig > 15 ? 15 : esteig);
            if (ltype == KSP_CHEBYSHEV) oksp_estimate_eigs(&m->sm[l]);
        } else {
            oksp_init(&m->sm[l], ctype, &m->A[l], cpc);
            m->sm[l].maxit = (ctype == KSP_PREONLY) ? 1 : oopts_int(o, "mg_coarse_ksp_max_it", 10000);
        }
        m->sm[l].guess_nonzero = 1;
        m->sm[l].normtype = NORM_NONE;
        if (l < L) {
            m->x[l] = (double *)calloc(n, sizeof(double));
            m->b[l] = (double *)calloc(n, sizeof(double));
        }
        m->r[l] = (double *)calloc(n, sizeof(double));
    }
    if (nlev == 1) m->sm[0].guess_nonzero = 0;
    return 0;
}

static void msmg_cycle(msmg *m, int l) {
    int reason; double r0, rn;
    if (l == 0) {
        if (m->sm[0].type == KSP_PREONLY) opc_apply(&m->sm[0].pc, m->b[0], m->x[0]);
        else oksp_solve(&m->sm[0], m->b[0], m->x[0], &reason, &r0, &rn);
        return;
    }
    const ocsr *P = &m->P[l];
    const long nf = m->g[l].n, nc = m->g[l - 1].n;
    oksp_solve(&m->sm[l], m->b[l], m->x[l], &reason, &r0, &rn);
    msmg_resid(m, l, m->b[l], m->x[l], m->r[l]);
    ovset(nc, 0.0, m->b[l - 1]);
    for (long f = 0; f < nf; f++)
        for (int64_t q = P->rp[f]; q < P->rp[f + 1]; q++) m->b[l - 1][P->ci[q]] += P->v[q] * m->r[l][f];
    ovset(nc, 0.0, m->x[l - 1]);
    for (int c = 0; c < m->cycles; c++) msmg_cycle(m, l - 1);
    for (long f = 0; f < nf; f++) {
        double s = 0.0;
        for (int64_t q = P->rp[f]; q < P->rp[f + 1]; q++) s += P->v[q] * m->x[l - 1][P->ci[q]];
        m->x[l][f] += s;
    }
    oksp_solve(&m->sm[l], m->b[l], m->x[l], &reason, &r0, &rn);
}

static void msmg_apply(void *ctx, const double *r, double *z) {
    msmg *m = (msmg *)ctx;
    const int L = m->nlev - 1;
    if (m->nlev == 1) {   /* a one-level PCMG applies the smoother from zero */
        int reason; double r0, rn;
        oksp_solve(&m->sm[0], r, z, &reason, &r0, &rn);
        return;
    }
    m->b[L] = (double *)r;
    m->x[L] = z;
    ovset(m->g[L].n, 0.0, z);
    msmg_cycle(m, L);
    m->b[L] = NULL; m->x[L] = NULL;
}

/* ------------------------------------------------------------- SNES */

enum { LS_BT = 0, LS_BASIC };

typedef struct {
    oopts *o;
    msctx  m;
    int    ktype, pctype, kmaxit, restart, ls, maxit, monitor, msmon, ksp_reason, sym, mg_levels;
    int    fd, mf;       /* -snes_fd_color: the colored Jacobian is the preconditioning matrix (else the
                            registered Poisson Jacobian); -snes_mf_operator: the operator is matrix-free */
    int    nranks;       /* simulated ranks (-oracle_nranks) */
    double krtol, katol, kdtol, rtol, atol, stol, divtol;
    char   symtol[32];
} ms_cfg;

static void petsc_g(char *b, int n, double v) {
    snprintf(b, n, "%g", v);
    if (!strpbrk(b, ".eEin")) strncat(b, ".", n - strlen(b) - 1);
}

static const char *snes_reason_name(int r) {
    switch (r) {
    case 2: return "CONVERGED_FNORM_ABS";
    case 3: return "CONVERGED_FNORM_RELATIVE";
    case 4: return "CONVERGED_SNORM_RELATIVE";
    case 5: return "CONVERGED_ITS";
    case -3: return "DIVERGED_LINEAR_SOLVE";
    case -4: return "DIVERGED_FNORM_NAN";
    case -5: return "DIVERGED_MAX_IT";
    case -6: return "DIVERGED_LINE_SEARCH";
    case -9: return "DIVERGED_DTOL";
    }
    return "CONVERGED_ITERATING";
}

static void snes_line(char *out, int outlen, int *pos, int ind, int it, double f) {
    if (f > 1e-9) oprintf(out, outlen, pos, "%*s%3d SNES Function norm %g\n", ind, "", it, f);
    else if (f > 1e-11) oprintf(out, outlen, pos, "%*s%3d SNES Function norm %5.3e\n", ind, "", it, f);
    else oprintf(out, outlen, pos, "%*s%3d SNES Function norm < 1.e-11\n", ind, "", it);
}

static int snes_converged(const ms_cfg *c, int it, double xnorm, double snorm, double fnorm, double *ttol,
                          double *rnorm0) {
    if (it == 0) { *ttol = fnorm * c->rtol; *rnorm0 = fnorm; }
    if (isnan(fnorm) || isinf(fnorm)) return -4;
    if (fnorm < c->atol) return 2;
    if (it) {
        if (fnorm <= *ttol) return 3;
        if (snorm < c->stol * xnorm) return 4;
        if (c->divtol > 0 && fnorm > c->divtol * *rnorm0) return -9;
    }
    return 0;
}

/* MSEMonitor (minimal.c:305-362): area by tensor Gauss-Legendre on Q1
 * cells, and the extreme diffusivities at the quadrature points */
static void ms_area(const ogrid *g, const msctx *m, const double *u, double *area, double *Dmin, double *Dmax) {
    static const double xi[3][3] = {{0.0, 0, 0}, {-0.577350269189626, 0.577350269189626, 0},
                                    {-0.774596669241483, 0.0, 0.774596669241483}};
    static const double wt[3][3] = {{2.0, 0, 0}, {1.0, 1.0, 0}, {0.555555555555556, 0.888888888888889, 0.555555555555556}};
    const int qd = m->quaddeg - 1, nq = m->quaddeg;
    const int mx = g->mx, my = g->my;
    const double hx = 1.0 / (mx - 1), hy = 1.0 / (my - 1);
    double a = 0.0, dmin = INFINITY, dmax = 0.0;
    for (int j = 1; j < my; j++) {
        const double yj = j * hy;
        for (int i = 1; i < mx; i++) {
            const double xi_ = i * hx;
            const double u11 = u[(long)j * mx + i], u01 = u[(long)j * mx + i - 1];
            const double u10 = u[(long)(j - 1) * mx + i], u00 = u[(long)(j - 1) * mx + i - 1];
            for (int r = 0; r < nq; r++) {
                const double x = xi_ - hx + hx * 0.5 * (xi[qd][r] + 1);
                for (int s = 0; s < nq; s++) {
                    const double y = yj - hy + hy * 0.5 * (xi[qd][s] + 1);
                    double ux = (u11 - u01) * (y - (yj - hy)) + (u10 - u00) * (yj - y);
                    ux /= hx * hy;
                    double uy = (u11 - u10) * (x - (xi_ - hx)) + (u01 - u00) * (xi_ - x);
                    uy /= hx * hy;
                    const double W = ux * ux + uy * uy;
                    const double D = ms_D(W, m->q);
                    if (D < dmin) dmin = D;
                    if (D > dmax) dmax = D;
                    a += wt[qd][r] * wt[qd][s] * sqrt(1.0 + W);
                }
            }
        }
    }
    a *= hx * hy / 4.0;
    *area = a; *Dmin = dmin; *Dmax = dmax;
}

typedef struct {
    int its, reason, last_kspits, ksp_total;
    double fnorm0, fnorm;
} ms_stage;

/* BT or basic line search: X <- X - lambda Y, F <- F(X) */
static int line_search(const ms_cfg *cfg, const ogrid *g, const ocsr *J, msmf *mf, double *X, double *F,
                       double *fnorm, const double *Y, double *xnorm, double *ynorm, double *W, double *G) {
    const long n = g->n;
    const double alpha = 1e-4, steptol = 1e-12;
    double lambda = 1.0;
    *ynorm = ovnorm2(n, Y);
    *xnorm = ovnorm2(n, X);
#define TRIAL(lam)                                                  \
    do {                                                            \
        for (long i = 0; i < n; i++) W[i] = X[i] - (lam) * Y[i];   \
        ms_residual(g, &cfg->m, W, G);                              \
        gnorm = ovnorm2(n, G);                                      \
    } while (0)
    double gnorm;
    if (cfg->ls == LS_BASIC) {
        TRIAL(1.0);
        memcpy(X, W, sizeof(double) * n); memcpy(F, G, sizeof(double) * n);
        *fnorm = gnorm;
        *xnorm = ovnorm2(n, X);
        return 0;
    }
    if (*ynorm == 0.0) return 1;
    if (mf) msmf_mult(mf, Y, W);   /* SNES's Jacobian is the matrix-free operator */
    else ocsr_mult(J, Y, W);
    double initslope = ovdot(n, F, W);
    if (initslope > 0.0) initslope = -initslope;
    if (initslope == 0.0) initslope = -1.0;
    const double f = (*fnorm) * (*fnorm);
    TRIAL(lambda);
    double g2 = gnorm * gnorm;
    int ok = 0;
    if (0.5 * g2 <= 0.5 * f + lambda * alpha * initslope) ok = 1;
    else {
        double rellength = 0.0;
        for (long i = 0; i < n; i++) {
            const double t = fabs(Y[i]) / (fabs(X[i]) > 1.0 ? fabs(X[i]) : 1.0);
            if (t > rellength) rellength = t;
        }
        const double minlambda = steptol / rellength;
        double lambdatemp = -initslope / (g2 - f - 2.0 * lambda * initslope);
        double lambdaprev = lambda, gprev = g2;
        if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
        lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
        TRIAL(lambda);
        g2 = gnorm * gnorm;
        if (0.5 * g2 < 0.5 * f + lambda * alpha * initslope) ok = 1;
        while (!ok) {
            if (lambda <= minlambda) break;
            const double t1 = 0.5 * (g2 - f) - lambda * initslope;
            const double t2 = 0.5 * (gprev - f) - lambdaprev * initslope;
            const double a = (t1 / (lambda * lambda) - t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
            const double b = (-lambdaprev * t1 / (lambda * lambda) + lambda * t2 / (lambdaprev * lambdaprev)) /
                             (lambda - lambdaprev);
            double d = b * b - 3.0 * a * initslope;
            if (d < 0.0) d = 0.0;
            lambdatemp = (a == 0.0) ? -initslope / (2.0 * b) : (-b + sqrt(d)) / (3.0 * a);
            lambdaprev = lambda; gprev = g2;
            if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
            lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
            TRIAL(lambda);
            g2 = gnorm * gnorm;
            if (0.5 * g2 <= 0.5 * f + lambda * alpha * initslope) ok = 1;
        }
    }
#undef TRIAL
    memcpy(X, W, sizeof(double) * n); memcpy(F, G, sizeof(double) * n);
    *fnorm = gnorm;
    *xnorm = ovnorm2(n, X);
    *ynorm *= lambda;
    return ok ? 0 : 1;
}

/* SNESSolve_NEWTONLS on grid g from X (in/out) */
static int newtonls(const ms_cfg *cfg, const ogrid *g, double *X, int ind, ms_stage *st, char *out, int outlen,
                    int *pos, char *err, int errlen) {
    const long n = g->n;
    memset(st, 0, sizeof *st);
    double *F = (double *)malloc(sizeof(double) * n), *Y = (double *)calloc(n, sizeof(double));
    double *W = (double *)malloc(sizeof(double) * n), *G = (double *)malloc(sizeof(double) * n);
    double *Wmf = (double *)malloc(sizeof(double) * n);
    int e = 0;
    if (cfg->sym) oprintf(out, outlen, pos, "Matrix is symmetric (tolerance %s)\n", cfg->symtol);
    ms_residual(g, &cfg->m, X, F);
    double fnorm = ovnorm2(n, F), xnorm = ovnorm2(n, X), ynorm = 0.0, ttol = 0.0, rnorm0 = 0.0;
    st->fnorm0 = st->fnorm = fnorm;
    if (cfg->monitor) snes_line(out, outlen, pos, ind, 0, fnorm);
    if (cfg->msmon) {
        double a, d0, d1;
        ms_area(g, &cfg->m, X, &a, &d0, &d1);
        oprintf(out, outlen, pos, "%*sarea = %.8f; %.4f <= D <= %.4f\n", ind, "", a, d0, d1);
    }
    int reason = snes_converged(cfg, 0, 0.0, 0.0, fnorm, &ttol, &rnorm0);
    int i = 0, iter = 0;
    for (; !reason && i < cfg->maxit; i++) {
        ocsr J;
        msmf mfc = {g, &cfg->m, X, F, Wmf, 0};
        if (cfg->fd) ms_fdjac(g, &cfg->m, X, F, &J);   /* the preconditioning matrix */
        else ms_poisson(g, &J);
        if (cfg->sym) {
            const int symm = ocsr_is_symmetric(&J, atof(cfg->symtol));
            oprintf(out, outlen, pos, "Matrix is %ssymmetric (tolerance %s)\n", symm ? "" : "not ", cfg->symtol);
        }
        msmg mg;
        int pct = cfg->pctype;
        if (pct == PC_MG) {
            const int nlev = cfg->mg_levels;
            e = msmg_setup(&mg, g, &cfg->m, nlev, X, &J, cfg->fd, cfg->mf ? &mfc : NULL, cfg->nranks, cfg->o, err,
                           errlen);
            if (e) { ocsr_free(&J); break; }
            pct = PC_SHELL;
        }
        oksp k;
        oksp_init(&k, cfg->ktype, &J, pct);
        if (pct == PC_SHELL) { k.pc.shell = msmg_apply; k.pc.shell_ctx = &mg; }
        if (cfg->mf) { k.amult = msmf_mult; k.actx = &mfc; }
        k.rtol = cfg->krtol; k.atol = cfg->katol; k.dtol = cfg->kdtol; k.maxit = cfg->kmaxit;
        k.restart = cfg->restart;
        int kreason = 0;
        double r0, rn;
        const int kits = oksp_solve(&k, F, Y, &kreason, &r0, &rn);
        oksp_free(&k);
        if (cfg->pctype == PC_MG) msmg_free(&mg);
        st->last_kspits = kits;
        st->ksp_total += kits;
        if (cfg->ksp_reason) {
            if (kreason > 0)
                oprintf(out, outlen, pos, "%*s    Linear solve converged due to %s iterations %d\n", ind, "",
                        oracle_reason_name(kreason), kits);
            else
                oprintf(out, outlen, pos, "%*s    Linear solve did not converge due to %s iterations %d\n", ind, "",
                        oracle_reason_name(kreason), kits);
        }
        if (kreason < 0) { ocsr_free(&J); reason = -3; break; }
        double gnorm = fnorm;
        const int lsfail = line_search(cfg, g, &J, cfg->mf ? &mfc : NULL, X, F, &gnorm, Y, &xnorm, &ynorm, W, G);
        ocsr_free(&J);
        if (lsfail) {
            reason = (cfg->stol * xnorm > ynorm) ? 4 : -6;
            break;
        }
        fnorm = gnorm;
        iter = i + 1;
        if (cfg->monitor) snes_line(out, outlen, pos, ind, iter, fnorm);
        if (cfg->msmon) {
            double a, d0, d1;
            ms_area(g, &cfg->m, X, &a, &d0, &d1);
            oprintf(out, outlen, pos, "%*sarea = %.8f; %.4f <= D <= %.4f\n", ind, "", a, d0, d1);
        }
        reason = snes_converged(cfg, iter, xnorm, ynorm, fnorm, &ttol, &rnorm0);
    }
    if (!reason && !e && i >= cfg->maxit) reason = -5;
    st->its = iter;
    st->reason = reason;
    st->fnorm = fnorm;
    free(F); free(Y); free(W); free(G); free(Wmf);
    return e;
}

/* exported for the tests: the residual and the colored finite-difference
 * Jacobian (9 coefficients per row, offsets (ox, oy) in the order
 * oy = -1..1 outer, ox = -1..1 inner; zero outside the grid) on an mx x my
 * grid for problem 0 tent / 1 catenoid with exponent q, H and c */
void oracle_minimal_residual(int mx, int my, int problem, double q, double H, double c, const double *u, double *F) {
    msctx m = {problem, 3, q, H, c};
    ogrid g;
    ogrid_set(&g, 2, mx, my, 1, 1.0, 1.0, 1.0);
    ms_residual(&g, &m, u, F);
}

void oracle_minimal_fdjac(int mx, int my, int problem, double q, double H, double c, const double *u, double *J9) {
    msctx m = {problem, 3, q, H, c};
    ogrid g;
    ogrid_set(&g, 2, mx, my, 1, 1.0, 1.0, 1.0);
    double *F = (double *)malloc(sizeof(double) * g.n);
    ms_residual(&g, &m, u, F);
    ocsr J;
    ms_fdjac(&g, &m, u, F, &J);
    memset(J9, 0, sizeof(double) * 9 * g.n);
    for (long r = 0; r < g.n; r++)
        for (int64_t p = J.rp[r]; p < J.rp[r + 1]; p++) {
            const long col = J.ci[p];
            const int ox = (int)(col % mx) - (int)(r % mx), oy = (int)(col / mx) - (int)(r / mx);
            J9[r * 9 + (oy + 1) * 3 + (ox + 1)] = J.v[p];
        }
    ocsr_free(&J);
    free(F);
}

void oracle_minimal_mffd(int mx, int my, int problem, double q, double H, double c, const double *u,
                         const double *a, double *y) {
    msctx m = {problem, 3, q, H, c};
    ogrid g;
    ogrid_set(&g, 2, mx, my, 1, 1.0, 1.0, 1.0);
    double *F = (double *)malloc(sizeof(double) * g.n), *w = (double *)malloc(sizeof(double) * g.n);
    ms_residual(&g, &m, u, F);
    msmf mf = {&g, &m, u, F, w, 0};
    msmf_mult(&mf, a, y);
    free(F); free(w);
}

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int oracle_minimal_run(int argc, const char **argv, minimal_result *res, double *u_out, char *out, int outlen) {
    minimal_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    if (out && outlen > 0) out[0] = 0;
    int pos = 0;
    char err[512] = "";
    oopts o;
    oopts_parse(&o, argc, argv);
    ms_cfg c;
    memset(&c, 0, sizeof c);
    c.o = &o;
    /* minimal.c:70-104: the ms_ options */
    c.m.q = oopts_real(&o, "ms_q", -0.5);
    c.m.H = oopts_real(&o, "ms_tent_H", 1.0);
    c.m.c = oopts_real(&o, "ms_catenoid_c", 1.1);
    c.m.quaddeg = oopts_int(&o, "ms_quaddegree", 3);
    const int exact_init = oopts_bool(&o, "ms_exact_init", 0);
    c.msmon = oopts_bool(&o, "ms_monitor", 0);
    const char *pb = oopts_get(&o, "ms_problem");
    c.m.problem = MS_CATENOID;
    if (pb) {
        if (!strcmp(pb, "tent")) c.m.problem = MS_TENT;
        else if (strcmp(pb, "catenoid")) {
            oprintf(out, outlen, &pos, "[0]PETSC ERROR: unknown -ms_problem %s\n", pb);
            return 62;
        }
    }
    if (c.m.quaddeg < 1 || c.m.quaddeg > 3) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: -ms_quaddegree must be 1, 2 or 3\n");
        return 63;
    }
    /* minimal.c:107-129 */
    if (c.m.problem == MS_TENT && exact_init) {
        oprintf(out, outlen, &pos,
                "[0]PETSC ERROR: initialization with exact solution only possible for -mse_problem catenoid\n");
        return 2;
    }
    if (c.m.problem == MS_CATENOID && c.m.c < 1.0) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: catenoid exact solution only valid if c >= 1\n");
        return 3;
    }
    if (c.m.problem == MS_CATENOID && exact_init && c.m.q != -0.5) {
        oprintf(out, outlen, &pos,
                "[0]PETSC ERROR: initialization with catenoid exact solution only possible if q=-0.5\n");
        return 4;
    }
    const char *st = oopts_get(&o, "snes_type");
    if ((st && strcmp(st, "newtonls")) || oopts_has(&o, "snes_mf")) {
        oprintf(out, outlen, &pos,
                "[0]PETSC ERROR: oracle: only -snes_type newtonls with -snes_fd_color or -snes_mf_operator\n");
        return 56;
    }
    c.fd = oopts_has(&o, "snes_fd_color");
    c.mf = oopts_has(&o, "snes_mf_operator");
    c.nranks = oopts_int(&o, "oracle_nranks", 1);
    if (c.nranks < 1 || c.nranks > 16) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: -oracle_nranks must be 1..16\n");
        return 63;
    }
    const char *lt = oopts_get(&o, "snes_linesearch_type");
    c.ls = (lt && !strcmp(lt, "basic")) ? LS_BASIC : LS_BT;
    if (lt && strcmp(lt, "basic") && strcmp(lt, "bt")) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: -snes_linesearch_type %s not supported\n", lt);
        return 56;
    }
    c.ktype = oracle_parse_ksp(oopts_get(&o, "ksp_type"), KSP_GMRES);
    c.pctype = oracle_parse_pc(oopts_get(&o, "pc_type"), PC_ILU);
    if (c.ktype < 0 || c.pctype < 0 || c.ktype == KSP_PREONLY) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: unsupported -ksp_type/-pc_type\n");
        return 56;
    }
    c.krtol = oopts_real(&o, "ksp_rtol", 1e-5);
    c.katol = oopts_real(&o, "ksp_atol", 1e-50);
    c.kdtol = oopts_real(&o, "ksp_divtol", 1e5);
    c.kmaxit = oopts_int(&o, "ksp_max_it", 10000);
    c.restart = oopts_int(&o, "ksp_gmres_restart", 30);
    c.rtol = oopts_real(&o, "snes_rtol", 1e-8);
    c.atol = oopts_real(&o, "snes_atol", 1e-50);
    c.stol = oopts_real(&o, "snes_stol", 1e-8);
    c.divtol = oopts_real(&o, "snes_divergence_tolerance", 1e4);
    c.maxit = oopts_int(&o, "snes_max_it", 50);
    c.monitor = oopts_has(&o, "snes_monitor_short") || oopts_has(&o, "snes_monitor");
    c.ksp_reason = oopts_has(&o, "ksp_converged_reason");
    const char *sv = oopts_get(&o, "mat_is_symmetric");
    c.sym = sv != NULL;
    if (c.sym) petsc_g(c.symtol, sizeof c.symtol, *sv ? atof(sv) : 0.0);
    const int snes_reason = oopts_has(&o, "snes_converged_reason");
    /* grid (minimal.c:141-148): 3x3 base, -da_grid_x/y, -da_refine */
    int mx = oopts_int(&o, "da_grid_x", 3), my = oopts_int(&o, "da_grid_y", 3);
    const int refine = oopts_int(&o, "da_refine", 0), seq = oopts_int(&o, "snes_grid_sequence", 0);
    if (seq < 0 || seq > 12 || mx < 2 || my < 2) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: grid options out of range\n");
        return 63;
    }
    for (int r = 0; r < refine; r++) { mx = 2 * (mx - 1) + 1; my = 2 * (my - 1) + 1; }
    const int lev_opt = oopts_int(&o, "pc_mg_levels", -1);
    ogrid g;
    ogrid_set(&g, 2, mx, my, 1, 1.0, 1.0, 1.0);
    double *X = (double *)calloc(g.n, sizeof(double));
    {   /* InitialState(ZEROS, g on the boundary) or FormExactFromG */
        const double hx = 1.0 / (mx - 1), hy = 1.0 / (my - 1);
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++)
                if (exact_init || i == 0 || j == 0 || i == mx - 1 || j == my - 1)
                    X[(long)j * mx + i] = ms_g(&c.m, i * hx, j * hy);
    }
    const double t0 = now_ms();
    ms_stage sg;
    int e = 0;
    res->stages = seq + 1;
    for (int stage = 0; stage <= seq; stage++) {
        c.mg_levels = lev_opt > 0 ? lev_opt : refine + stage + 1;
        const int ind = 2 * (seq - stage);
        e = newtonls(&c, &g, X, ind, &sg, out, outlen, &pos, err, sizeof err);
        if (e) break;
        if (stage < 16) res->stage_its[stage] = sg.its;
        res->ksp_its_total += sg.ksp_total;
        if (snes_reason) {
            if (sg.reason > 0)
                oprintf(out, outlen, &pos, "%*s  Nonlinear solve converged due to %s iterations %d\n", ind, "",
                        snes_reason_name(sg.reason), sg.its);
            else
                oprintf(out, outlen, &pos, "%*s  Nonlinear solve did not converge due to %s iterations %d\n", ind,
                        "", snes_reason_name(sg.reason), sg.its);
        }
        if (stage < seq) {   /* DMRefine + MatInterpolate of the solution */
            ogrid gf;
            ogrid_set(&gf, 2, 2 * (g.mx - 1) + 1, 2 * (g.my - 1) + 1, 1, 1.0, 1.0, 1.0);
            double *Xf = (double *)calloc(gf.n, sizeof(double));
            oq1_interp_add(&g, &gf, X, Xf);
            free(X);
            X = Xf;
            g = gf;
        }
    }
    res->solve_ms = now_ms() - t0;
    if (e) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: %s\n", err);
        free(X);
        return e;
    }
    /* report (minimal.c:184-204) */
    res->mx = g.mx; res->my = g.my;
    res->its = sg.its; res->reason = sg.reason; res->last_ksp_its = sg.last_kspits;
    res->fnorm0 = sg.fnorm0; res->fnorm = sg.fnorm;
    {
        double a, d0, d1;
        ms_area(&g, &c.m, X, &a, &d0, &d1);
        res->area = a;
    }
    oprintf(out, outlen, &pos, "done on %d x %d grid and problem %s", g.mx, g.my,
            c.m.problem == MS_TENT ? "tent" : "catenoid");
    if (c.m.problem == MS_CATENOID && c.m.q == -0.5) {
        const double hx = 1.0 / (g.mx - 1), hy = 1.0 / (g.my - 1);
        double emax = 0.0;
        for (int j = 0; j < g.my; j++)
            for (int i = 0; i < g.mx; i++) {
                const double d = fabs(X[(long)j * g.mx + i] - ms_g(&c.m, i * hx, j * hy));
                if (d > emax) emax = d;
            }
        res->errinf = emax;
        oprintf(out, outlen, &pos, ":  error |u-uexact|_inf = %.5e\n", emax);
    } else {
        oprintf(out, outlen, &pos, " ...\n");
    }
    if (u_out) memcpy(u_out, X, sizeof(double) * g.n);
    free(X);
    return 0;
}
