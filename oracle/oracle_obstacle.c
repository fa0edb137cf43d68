/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * c/ch12/obstacle.c restated: the obstacle problem -Lap u = 0, u >= psi on
 * (-2,2)^2 with u = u_exact on the boundary, solved as a variational
 * inequality by PETSc's reduced-space Newton method (SNESVINEWTONRSLS,
 * obstacle.c:118-121) over the ch6 Poisson residual and Jacobian
 * (obstacle.c:123-127, c/ch6/poissonfunctions.c:29-65,152-193), with
 * optional grid sequencing (-snes_grid_sequence, c/ch12/makefile:23-27).
 *
 * PETSc semantics restated (PETSc is not vendored; no version is pinned):
 *   - SNESSolve_VINEWTONRSLS: project the iterate onto the bounds; F(x);
 *     fnorm = |F| over the inactive set; per iteration the active set
 *     {x <= xl + 1e-8 and F > 0} or {x >= xu - 1e-8 and F < 0}, the
 *     inactive-set submatrix J_II, KSPSolve(J_II, F_I) for the direction
 *     (zero on the active set), the line search, SNESConvergedDefault;
 *   - SNESLineSearchApply_BT with alpha = 0 (the RSLS default): trial
 *     points projected onto the bounds, the inactive-set norm, quadratic
 *     then cubic backtracking (Dennis & Schnabel, A6.3.1);
 *   - SNESSolve_VINEWTONSSLS (-snes_type vinewtonssls): the Fischer-
 *     Burmeister reformulation phi_i = Phi(x_i - l_i, F_i), Phi(a, b) = a + b
 *     - sqrt(a^2 + b^2) (the obstacle sets lower bounds only), |phi| as the
 *     SNES norm; per iteration the semismooth Jacobian diag(Db) J + diag(Da)
 *     from the B-subdifferential (Da_i = Phi_a(x_i - l_i, F_i), Db_i =
 *     Phi_a(F_i, x_i - l_i), Phi_a(0, 0) taken as 1 - sqrt(2)/2), KSPSolve
 *     for the direction, the BT line search on phi with trial points
 *     projected onto the bounds (alpha 0; test4 does not tell alpha 0 from
 *     1e-4), no projection of the initial iterate;
 *   - PCMG under RSLS uses Galerkin coarse operators on coarse inactive
 *     sets (a coarse vertex is inactive when the restriction of the fine
 *     inactive indicator is nonzero there) with the interpolation
 *     restricted to the inactive rows and columns.
 *
 * Pinning: obstacle.test1 (|F| history 3.86571, 1.3323, < 1e-11, 11 CG/ILU
 * iterations per Newton step, the error line) and obstacle.test3 (grid-
 * sequenced SNES iteration counts 1, 2, 2, 3, last KSP count 4 and the
 * error line), obstacle.test4 (SSLS, grid-sequenced Newton counts 4, 6, 5,
 * last KSP count 6 and the error line; unprojected trial points give other
 * counts).  A one-level PCMG (test3's 3x3 grid) applies the level
 * smoother, as PETSc's PCMGSetLevels makes it.  The backtracking branch of
 * the line search is not reached by those tests: parity unpinned there.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"
#include "oracle_internal.h"

/* --------------------------------------------- problem (obstacle.c:15-69) */

static double obs_psi(double x, double y) {
    const double r = sqrt(x * x + y * y), r0 = 0.9, psi0 = sqrt(1.0 - r0 * r0), dpsi0 = -r0 / psi0;
    return r <= r0 ? sqrt(1.0 - r * r) : psi0 + dpsi0 * (r - r0);
}

static double obs_uexact(double x, double y) {
    const double afree = 0.697965148223374, A = 0.680259411891719, B = 0.471519893402112;
    const double r = sqrt(x * x + y * y);
    return r <= afree ? obs_psi(x, y) : -A * log(r) + B;
}

#define OBS_AEXACT 0.697965148223374

/* Poisson2DFunctionLocal (poissonfunctions.c:29-65) on (-2,2)^2, cx = cy = 1,
 * g = u_exact, f = 0 */
static void obs_residual(const ogrid *g, const double *u, double *F) {
    const int mx = g->mx, my = g->my;
    const double hx = g->hx, hy = g->hy, scx = hy / hx, scy = hx / hy, scdiag = 2.0 * (scx + scy);
    for (int j = 0; j < my; j++) {
        const double y = -2.0 + j * hy;
        for (int i = 0; i < mx; i++) {
            const double x = -2.0 + i * hx;
            const long q = (long)j * mx + i;
            if (i == 0 || i == mx - 1 || j == 0 || j == my - 1) {
                F[q] = scdiag * (u[q] - obs_uexact(x, y));
                continue;
            }
            const double ue = (i + 1 == mx - 1) ? obs_uexact(x + hx, y) : u[q + 1];
            const double uw = (i - 1 == 0) ? obs_uexact(x - hx, y) : u[q - 1];
            const double un = (j + 1 == my - 1) ? obs_uexact(x, y + hy) : u[q + mx];
            const double us = (j - 1 == 0) ? obs_uexact(x, y - hy) : u[q - mx];
            F[q] = scdiag * u[q] - scx * (uw + ue) - scy * (us + un);
        }
    }
}

/* FormBounds (obstacle.c:234-253) */
static void obs_bounds(const ogrid *g, double *xl) {
    for (int j = 0; j < g->my; j++)
        for (int i = 0; i < g->mx; i++) xl[(long)j * g->mx + i] = obs_psi(-2.0 + i * g->hx, -2.0 + j * g->hy);
}

/* ------------------------------------------------------- VI set helpers */

#define VI_ZEROTOL 1.0e-8

/* SNESVIGetActiveSetIS / SNESVIComputeInactiveSetFnorm predicate */
static int vi_active(double x, double f, double xl, double xu) {
    return (x <= xl + VI_ZEROTOL && f > 0.0) || (x >= xu - VI_ZEROTOL && f < 0.0);
}

static double vi_fnorm(long n, const double *F, const double *X, const double *xl, const double *xu) {
    double s = 0.0;
    for (long i = 0; i < n; i++)
        if (!vi_active(X[i], F[i], xl[i], xu ? xu[i] : INFINITY)) s += F[i] * F[i];
    return sqrt(s);
}

/* SNESVIProjectOntoBounds */
static void vi_project(long n, double *X, const double *xl, const double *xu) {
    for (long i = 0; i < n; i++) {
        if (X[i] < xl[i]) X[i] = xl[i];
        else if (xu && X[i] > xu[i]) X[i] = xu[i];
    }
}

/* rows and columns `keep` of A (a principal submatrix), renumbered */
static void csr_submatrix(const ocsr *A, long nk, const long *keep, const long *map, ocsr *S) {
    int64_t nnz = 0;
    for (long r = 0; r < nk; r++)
        for (int64_t q = A->rp[keep[r]]; q < A->rp[keep[r] + 1]; q++)
            if (map[A->ci[q]] >= 0) nnz++;
    S->n = nk;
    S->rp = (int64_t *)malloc(sizeof(int64_t) * (nk + 1));
    S->ci = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    S->v = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t w = 0;
    S->rp[0] = 0;
    for (long r = 0; r < nk; r++) {
        for (int64_t q = A->rp[keep[r]]; q < A->rp[keep[r] + 1]; q++)
            if (map[A->ci[q]] >= 0) { S->ci[w] = (int32_t)map[A->ci[q]]; S->v[w++] = A->v[q]; }
        S->rp[r + 1] = w;
    }
}

/* ------------------------------------------------ PCASM (restricted) */

/* The PETSc DMDA 2D process grid for P ranks of an mx x my grid
 * (DMSetUp_DA_2D's squarish choice) and its ownership ranges: rank
 * (rx, ry) = ry px + rx owns columns xs[rx] .. xs[rx+1]-1, rows ys[ry] .. */
static void dmda_split2(int mx, int my, int P, int *px, int *py, int xs[17], int ys[17]) {
    int m = (int)(0.5 + sqrt((double)mx * P / my));
    if (!m) m = 1;
    int n = 1;
    while (m > 0) {
        n = P / m;
        if (m * n == P) break;
        m--;
    }
    if (mx > my && m < n) { const int t = m; m = n; n = t; }
    *px = m; *py = n;
    xs[0] = 0;
    for (int q = 0; q < m; q++) xs[q + 1] = xs[q] + mx / m + (mx % m > q);
    ys[0] = 0;
    for (int q = 0; q < n; q++) ys[q + 1] = ys[q] + my / n + (my % n > q);
}

/* PCASM with one subdomain per simulated rank: the rank's rows of the
 * (reduced) matrix, grown by `overlap` levels of the matrix graph
 * (MatIncreaseOverlap), each solved exactly (-sub_pc_type lu), and the
 * restricted form (PC_ASM_RESTRICT, the default): a subdomain's solution is
 * added back on the rank's own rows only. */
typedef struct {
    int nq;
    long n;
    long *cnt, **idx;        /* subdomain q: cnt[q] indices, ascending */
    char **own;              /* whether idx[q][i] is owned by rank q */
    double **inv;            /* dense inverse of A[idx, idx], row-major */
    double *w1, *w2;
} oasm;

static void oasm_free(oasm *a) {
    for (int q = 0; q < a->nq; q++) { free(a->idx[q]); free(a->own[q]); free(a->inv[q]); }
    free(a->cnt); free(a->idx); free(a->own); free(a->inv); free(a->w1); free(a->w2);
    memset(a, 0, sizeof *a);
}

/* dense inverse by Gauss-Jordan with partial pivoting (an exact solve up to
 * rounding, as the subdomain LU is) */
static int dense_inverse(int m, double *A, double *X) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) X[(long)i * m + j] = (i == j) ? 1.0 : 0.0;
    for (int c = 0; c < m; c++) {
        int pr = c;
        for (int r = c + 1; r < m; r++)
            if (fabs(A[(long)r * m + c]) > fabs(A[(long)pr * m + c])) pr = r;
        if (A[(long)pr * m + c] == 0.0) return 1;
        if (pr != c)
            for (int j = 0; j < m; j++) {
                double t = A[(long)c * m + j]; A[(long)c * m + j] = A[(long)pr * m + j]; A[(long)pr * m + j] = t;
                t = X[(long)c * m + j]; X[(long)c * m + j] = X[(long)pr * m + j]; X[(long)pr * m + j] = t;
            }
        const double piv = A[(long)c * m + c];
        for (int j = 0; j < m; j++) { A[(long)c * m + j] /= piv; X[(long)c * m + j] /= piv; }
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            const double f = A[(long)r * m + c];
            if (f == 0.0) continue;
            for (int j = 0; j < m; j++) {
                A[(long)r * m + j] -= f * A[(long)c * m + j];
                X[(long)r * m + j] -= f * X[(long)c * m + j];
            }
        }
    }
    return 0;
}

/* S: the reduced matrix on the kept vertices keep[0..nk-1] of grid g */
static int oasm_setup(oasm *a, const ogrid *g, const ocsr *S, long nk, const long *keep, int P, int overlap,
                      char *err, int errlen) {
    memset(a, 0, sizeof *a);
    int px, py, xs[17], ys[17];
    dmda_split2(g->mx, g->my, P, &px, &py, xs, ys);
    a->nq = P; a->n = nk;
    a->cnt = (long *)calloc(P, sizeof(long));
    a->idx = (long **)calloc(P, sizeof(long *));
    a->own = (char **)calloc(P, sizeof(char *));
    a->inv = (double **)calloc(P, sizeof(double *));
    a->w1 = (double *)malloc(sizeof(double) * (nk ? nk : 1));
    a->w2 = (double *)malloc(sizeof(double) * (nk ? nk : 1));
    int *rank = (int *)malloc(sizeof(int) * (nk ? nk : 1));
    for (long r = 0; r < nk; r++) {
        const int i = (int)(keep[r] % g->mx), j = (int)(keep[r] / g->mx);
        int rx = 0, ry = 0;
        while (i >= xs[rx + 1]) rx++;
        while (j >= ys[ry + 1]) ry++;
        rank[r] = ry * px + rx;
    }
    char *in = (char *)malloc(nk ? nk : 1), *nx = (char *)malloc(nk ? nk : 1);
    for (int q = 0; q < P; q++) {
        for (long r = 0; r < nk; r++) in[r] = rank[r] == q;
        for (int o = 0; o < overlap; o++) {
            memcpy(nx, in, nk);
            for (long r = 0; r < nk; r++)
                if (in[r])
                    for (int64_t t = S->rp[r]; t < S->rp[r + 1]; t++) nx[S->ci[t]] = 1;
            memcpy(in, nx, nk);
        }
        long c = 0;
        for (long r = 0; r < nk; r++) c += in[r];
        a->cnt[q] = c;
        a->idx[q] = (long *)malloc(sizeof(long) * (c ? c : 1));
        a->own[q] = (char *)malloc(c ? c : 1);
        long w = 0;
        for (long r = 0; r < nk; r++)
            if (in[r]) { a->own[q][w] = rank[r] == q; a->idx[q][w++] = r; }
        if (c > 2048) {
            snprintf(err, errlen, "oracle: an ASM subdomain of %ld unknowns (dense solve up to 2048)", c);
            free(rank); free(in); free(nx);
            return 56;
        }
        if (!c) continue;
        double *A = (double *)calloc((size_t)c * c, sizeof(double));
        long *loc = (long *)malloc(sizeof(long) * nk);
        for (long r = 0; r < nk; r++) loc[r] = -1;
        for (long i2 = 0; i2 < c; i2++) loc[a->idx[q][i2]] = i2;
        for (long i2 = 0; i2 < c; i2++) {
            const long r = a->idx[q][i2];
            for (int64_t t = S->rp[r]; t < S->rp[r + 1]; t++)
                if (loc[S->ci[t]] >= 0) A[i2 * c + loc[S->ci[t]]] = S->v[t];
        }
        a->inv[q] = (double *)malloc(sizeof(double) * (size_t)c * c);
        const int bad = dense_inverse((int)c, A, a->inv[q]);
        free(A); free(loc);
        if (bad) {
            snprintf(err, errlen, "oracle: singular ASM subdomain matrix");
            free(rank); free(in); free(nx);
            return 76;
        }
    }
    free(rank); free(in); free(nx);
    return 0;
}

static void oasm_apply(void *ctx, const double *r, double *z) {
    oasm *a = (oasm *)ctx;
    for (long i = 0; i < a->n; i++) z[i] = 0.0;
    for (int q = 0; q < a->nq; q++) {
        const long c = a->cnt[q];
        for (long i = 0; i < c; i++) a->w1[i] = r[a->idx[q][i]];
        for (long i = 0; i < c; i++) {
            if (!a->own[q][i]) continue;
            double s2 = 0.0;
            for (long j = 0; j < c; j++) s2 += a->inv[q][i * c + j] * a->w1[j];
            z[a->idx[q][i]] += s2;
        }
    }
}

/* ----------------------------------------- PCMG on the inactive sets */

typedef struct {
    int     nlev, cycles;
    ogrid  *g;
    long   *n;          /* reduced size per level                     */
    long  **gidx;       /* reduced -> full vertex index per level      */
    ocsr   *A;          /* reduced operators; A[nlev-1] borrowed       */
    ocsr   *P;          /* P[l]: level l-1 -> l, rows n[l], cols n[l-1] */
    oksp   *sm;         /* smoother per level, coarse solver at 0      */
    double **x, **b, **r;
} vimg;

static void vimg_free(vimg *m) {
    for (int l = 0; l < m->nlev; l++) {
        oksp_free(&m->sm[l]);
        free(m->gidx[l]);
        if (l < m->nlev - 1) { ocsr_free(&m->A[l]); free(m->x[l]); free(m->b[l]); }
        if (m->P[l].v) ocsr_free(&m->P[l]);
        free(m->r[l]);
    }
    free(m->g); free(m->n); free(m->gidx); free(m->A); free(m->P); free(m->sm);
    free(m->x); free(m->b); free(m->r);
    memset(m, 0, sizeof *m);
}

/* DMCoarsen_SNESVI + DMCreateInterpolation_SNESVI + PCMG Galerkin: from the
 * finest inactive set (keep[0..nk-1], full indices) and J_II. */
static int vimg_setup(vimg *m, const ogrid *fine, int nlev, long nk, const long *keep, const ocsr *Aii, oopts *o,
                      char *err, int errlen) {
    memset(m, 0, sizeof *m);
    m->nlev = nlev;
    m->g = (ogrid *)calloc(nlev, sizeof(ogrid));
    m->n = (long *)calloc(nlev, sizeof(long));
    m->gidx = (long **)calloc(nlev, sizeof(long *));
    m->A = (ocsr *)calloc(nlev, sizeof(ocsr));
    m->P = (ocsr *)calloc(nlev, sizeof(ocsr));
    m->sm = (oksp *)calloc(nlev, sizeof(oksp));
    m->x = (double **)calloc(nlev, sizeof(double *));
    m->b = (double **)calloc(nlev, sizeof(double *));
    m->r = (double **)calloc(nlev, sizeof(double *));
    const char *ct = oopts_get(o, "pc_mg_cycle_type");
    m->cycles = (ct && !strcmp(ct, "w")) ? 2 : 1;
    const int L = nlev - 1;
    m->g[L] = *fine;
    for (int l = L; l > 0; l--)
        if (ogrid_coarsen(&m->g[l], &m->g[l - 1])) {
            snprintf(err, errlen, "oracle: grid of %d levels cannot be coarsened", nlev);
            return 63;
        }
    m->n[L] = nk;
    m->gidx[L] = (long *)malloc(sizeof(long) * (nk ? nk : 1));
    memcpy(m->gidx[L], keep, sizeof(long) * nk);
    m->A[L] = *Aii;
    for (int l = L; l > 0; l--) {
        const ogrid *gf = &m->g[l], *gc = &m->g[l - 1];
        ocsr Pf;
        oq1_build_P(gc, gf, &Pf);
        /* coarse inactive: P^T (fine inactive indicator) != 0 */
        double *mark = (double *)calloc(gc->n, sizeof(double));
        for (long r = 0; r < m->n[l]; r++) {
            const long f = m->gidx[l][r];
            for (int64_t q = Pf.rp[f]; q < Pf.rp[f + 1]; q++) mark[Pf.ci[q]] += Pf.v[q];
        }
        long *cmap = (long *)malloc(sizeof(long) * gc->n);
        long nc = 0;
        for (long c = 0; c < gc->n; c++) cmap[c] = mark[c] != 0.0 ? nc++ : -1;
        m->n[l - 1] = nc;
        m->gidx[l - 1] = (long *)malloc(sizeof(long) * (nc ? nc : 1));
        for (long c = 0; c < gc->n; c++)
            if (cmap[c] >= 0) m->gidx[l - 1][cmap[c]] = c;
        /* interpolation restricted to inactive fine rows and coarse columns */
        ocsr *P = &m->P[l];
        int64_t nnz = 0;
        for (long r = 0; r < m->n[l]; r++) nnz += Pf.rp[m->gidx[l][r] + 1] - Pf.rp[m->gidx[l][r]];
        P->n = m->n[l];
        P->rp = (int64_t *)malloc(sizeof(int64_t) * (P->n + 1));
        P->ci = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
        P->v = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
        int64_t w = 0;
        P->rp[0] = 0;
        for (long r = 0; r < m->n[l]; r++) {
            const long f = m->gidx[l][r];
            for (int64_t q = Pf.rp[f]; q < Pf.rp[f + 1]; q++)
                if (cmap[Pf.ci[q]] >= 0) { P->ci[w] = (int32_t)cmap[Pf.ci[q]]; P->v[w++] = Pf.v[q]; }
            P->rp[r + 1] = w;
        }
        oracle_galerkin(&m->A[l], P, nc, &m->A[l - 1]);
        free(mark); free(cmap);
        ocsr_free(&Pf);
    }
    /* smoothers and coarse solver: the PCMG defaults (Chebyshev, SOR, 2
     * iterations; coarse PREONLY + LU) and the -mg_levels_/-mg_coarse_ options */
    int ltype = oracle_parse_ksp(oopts_get(o, "mg_levels_ksp_type"), KSP_CHEBYSHEV);
    int lpc = oracle_parse_pc(oopts_get(o, "mg_levels_pc_type"), PC_SOR);
    int lits = oopts_int(o, "mg_levels_ksp_max_it", 2);
    double lscale = oopts_real(o, "mg_levels_ksp_richardson_scale", 1.0);
    int ctype = oracle_parse_ksp(oopts_get(o, "mg_coarse_ksp_type"), KSP_PREONLY);
    int cpc = oracle_parse_pc(oopts_get(o, "mg_coarse_pc_type"), PC_LU);
    int esteig = oopts_int(o, "mg_levels_ksp_chebyshev_esteig_steps", 10);
    if (ltype < 0 || lpc < 0 || lpc == PC_MG || ctype < 0 || cpc < 0 || cpc == PC_MG) {
        snprintf(err, errlen, "oracle: unsupported level smoother options");
        return 56;
    }
    for (int l = 0; l < nlev; l++) {
        const long n = m->n[l] ? m->n[l] : 1;
        if (l > 0 || nlev == 1) {   /* PCMGSetLevels: a single level is a smoother level */
            oksp_init(&m->sm[l], ltype, &m->A[l], lpc);
            m->sm[l].maxit = lits;
            m->sm[l].rich_scale = lscale;
            m->sm[l].esteig_steps = esteig < 1 ? 1 : (esteig > 15 ? 15 : esteig);
            m->sm[l].gidx = m->gidx[l];
            if (ltype == KSP_CHEBYSHEV && m->n[l] > 0) oksp_estimate_eigs(&m->sm[l]);
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

static void vimg_cycle(vimg *m, int l) {
    int reason; double r0, rn;
    if (l == 0) {
        if (m->n[0] == 0) return;
        if (m->sm[0].type == KSP_PREONLY) opc_apply(&m->sm[0].pc, m->b[0], m->x[0]);
        else oksp_solve(&m->sm[0], m->b[0], m->x[0], &reason, &r0, &rn);
        return;
    }
    const ocsr *P = &m->P[l];
    if (m->n[l] > 0) {
        oksp_solve(&m->sm[l], m->b[l], m->x[l], &reason, &r0, &rn);
        ocsr_resid(&m->A[l], m->b[l], m->x[l], m->r[l]);
    }
    /* b_c = P^T r */
    ovset(m->n[l - 1], 0.0, m->b[l - 1]);
    for (long f = 0; f < m->n[l]; f++)
        for (int64_t q = P->rp[f]; q < P->rp[f + 1]; q++) m->b[l - 1][P->ci[q]] += P->v[q] * m->r[l][f];
    ovset(m->n[l - 1], 0.0, m->x[l - 1]);
    for (int c = 0; c < m->cycles; c++) vimg_cycle(m, l - 1);
    for (long f = 0; f < m->n[l]; f++) {
        double s = 0.0;
        for (int64_t q = P->rp[f]; q < P->rp[f + 1]; q++) s += P->v[q] * m->x[l - 1][P->ci[q]];
        m->x[l][f] += s;
    }
    if (m->n[l] > 0) oksp_solve(&m->sm[l], m->b[l], m->x[l], &reason, &r0, &rn);
}

static void vimg_apply(void *ctx, const double *r, double *z) {
    vimg *m = (vimg *)ctx;
    const int L = m->nlev - 1;
    m->b[L] = (double *)r;
    m->x[L] = z;
    ovset(m->n[L], 0.0, z);
    vimg_cycle(m, L);
    m->b[L] = NULL; m->x[L] = NULL;
}

/* ------------------------------------------------- SNES configuration */

enum { LS_BT = 0, LS_BASIC };

typedef struct {
    oopts *o;
    int    ktype, pctype, normtype, kmaxit, knormset;
    double krtol, katol, kdtol;
    double rtol, atol, stol, divtol;
    int    maxit, ls;
    int    monitor, ksp_reason, sym, mg_levels;
    int    ssls;          /* -snes_type vinewtonssls */
    int    pcasm, nranks, asm_overlap;   /* -pc_type asm over -oracle_nranks simulated ranks */
    double ls_alpha;      /* the BT line search's sufficient-decrease constant */
    int    ls_project;    /* trial points projected onto the bounds */
    char   symtol[32];
} obs_cfg;

/* PetscPrintf's %g: an integral value keeps a trailing '.' ("0.") */
static void petsc_g(char *b, int n, double v) {
    snprintf(b, n, "%g", v);
    if (!strpbrk(b, ".eEin")) strncat(b, ".", n - strlen(b) - 1);
}

/* SNESConvergedReason names */
static const char *snes_reason_name(int r) {
    switch (r) {
    case 2: return "CONVERGED_FNORM_ABS";
    case 3: return "CONVERGED_FNORM_RELATIVE";
    case 4: return "CONVERGED_SNORM_RELATIVE";
    case 5: return "CONVERGED_ITS";
    case -1: return "DIVERGED_FUNCTION_DOMAIN";
    case -2: return "DIVERGED_FUNCTION_COUNT";
    case -3: return "DIVERGED_LINEAR_SOLVE";
    case -4: return "DIVERGED_FNORM_NAN";
    case -5: return "DIVERGED_MAX_IT";
    case -6: return "DIVERGED_LINE_SEARCH";
    case -8: return "DIVERGED_LOCAL_MIN";
    case -9: return "DIVERGED_DTOL";
    }
    return "CONVERGED_ITERATING";
}

/* SNESMonitorDefaultShort */
static void snes_line(char *out, int outlen, int *pos, int ind, int it, double f) {
    if (f > 1e-9) oprintf(out, outlen, pos, "%*s%3d SNES Function norm %g\n", ind, "", it, f);
    else if (f > 1e-11) oprintf(out, outlen, pos, "%*s%3d SNES Function norm %5.3e\n", ind, "", it, f);
    else oprintf(out, outlen, pos, "%*s%3d SNES Function norm < 1.e-11\n", ind, "", it);
}

/* SNESConvergedDefault */
static int snes_converged(const obs_cfg *c, int it, double xnorm, double snorm, double fnorm, double *ttol,
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

typedef struct {
    const ogrid *g;
    const ocsr  *A;
    const double *xl, *xu;
    double *W, *G;
    int ssls, project;   /* SSLS: G is the Fischer-Burmeister function, FW the residual at W */
    double *FW;
} lsctx;

/* the Fischer-Burmeister function and its derivative in its first argument
 * (the B-subdifferential element PETSc's SNESVINEWTONSSLS takes at 0) */
static double fb_phi(double a, double b) { return a + b - sqrt(a * a + b * b); }
static double fb_dphi(double a, double b) {
    if (fabs(a) <= 1.0e-30 && fabs(b) <= 1.0e-30) return 1.0 - sqrt(2.0) / 2.0;
    return 1.0 - a / sqrt(a * a + b * b);
}

/* SNESVIComputeFunction with lower bounds only: F = F(X), phi_i = Phi(x_i - l_i, F_i) */
static void ssls_function(const ogrid *g, const double *X, const double *xl, double *F, double *phi) {
    obs_residual(g, X, F);
    for (long i = 0; i < g->n; i++) phi[i] = fb_phi(X[i] - xl[i], F[i]);
}

/* trial point W = proj(X - lambda Y), G = F(W), returns |G| on W's inactive
 * set; SSLS: W = X - lambda Y (projected when asked), G = phi(W), |G| */
static double ls_trial(lsctx *c, const double *X, const double *Y, double lambda) {
    const long n = c->g->n;
    for (long i = 0; i < n; i++) c->W[i] = X[i] - lambda * Y[i];
    if (c->ssls) {
        if (c->project) vi_project(n, c->W, c->xl, c->xu);
        ssls_function(c->g, c->W, c->xl, c->FW, c->G);
        return ovnorm2(n, c->G);
    }
    vi_project(n, c->W, c->xl, c->xu);
    obs_residual(c->g, c->W, c->G);
    return vi_fnorm(n, c->G, c->W, c->xl, c->xu);
}

/* SNESLineSearchApply_BT (alpha = 0) or _Basic; returns 0 on success.  On
 * return X, F, *fnorm hold the accepted point; *ynorm = lambda |Y|. */
static int line_search(const obs_cfg *cfg, lsctx *c, double *X, double *F, double *fnorm, const double *Y,
                       double *xnorm, double *ynorm) {
    const long n = c->g->n;
    const double alpha = cfg->ls_alpha, steptol = 1e-12;
    double lambda = 1.0;
    *ynorm = ovnorm2(n, Y);
    *xnorm = ovnorm2(n, X);
    if (cfg->ls == LS_BASIC) {
        *fnorm = ls_trial(c, X, Y, 1.0);
        ovcopy(n, c->W, X); ovcopy(n, c->G, F);
        *xnorm = ovnorm2(n, X);
        return 0;
    }
    if (*ynorm == 0.0) return 1;
    /* initial slope of 1/2 |F|^2 along -Y: -(F, J Y) */
    ocsr_mult(c->A, Y, c->W);
    double initslope = ovdot(n, F, c->W);
    if (initslope > 0.0) initslope = -initslope;
    if (initslope == 0.0) initslope = -1.0;
    const double f = (*fnorm) * (*fnorm);
    double gnorm = ls_trial(c, X, Y, lambda), g = gnorm * gnorm;
    int ok = 0;
    if (0.5 * g <= 0.5 * f + lambda * alpha * initslope) ok = 1;
    else {
        double rellength = 0.0;
        for (long i = 0; i < n; i++) {
            const double t = fabs(Y[i]) / (fabs(X[i]) > 1.0 ? fabs(X[i]) : 1.0);
            if (t > rellength) rellength = t;
        }
        const double minlambda = steptol / rellength;
        /* quadratic model */
        double lambdatemp = -initslope / (g - f - 2.0 * lambda * initslope);
        double lambdaprev = lambda, gprev = g;
        if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
        lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
        gnorm = ls_trial(c, X, Y, lambda); g = gnorm * gnorm;
        if (0.5 * g < 0.5 * f + lambda * alpha * initslope) ok = 1;
        /* cubic model */
        while (!ok) {
            if (lambda <= minlambda) break;
            const double t1 = 0.5 * (g - f) - lambda * initslope;
            const double t2 = 0.5 * (gprev - f) - lambdaprev * initslope;
            const double a = (t1 / (lambda * lambda) - t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
            const double b = (-lambdaprev * t1 / (lambda * lambda) + lambda * t2 / (lambdaprev * lambdaprev)) /
                             (lambda - lambdaprev);
            double d = b * b - 3.0 * a * initslope;
            if (d < 0.0) d = 0.0;
            lambdatemp = (a == 0.0) ? -initslope / (2.0 * b) : (-b + sqrt(d)) / (3.0 * a);
            lambdaprev = lambda; gprev = g;
            if (lambdatemp > 0.5 * lambda) lambdatemp = 0.5 * lambda;
            lambda = lambdatemp <= 0.1 * lambda ? 0.1 * lambda : lambdatemp;
            gnorm = ls_trial(c, X, Y, lambda); g = gnorm * gnorm;
            if (0.5 * g <= 0.5 * f + lambda * alpha * initslope) ok = 1;
        }
    }
    ovcopy(n, c->W, X); ovcopy(n, c->G, F);
    *fnorm = gnorm;
    *xnorm = ovnorm2(n, X);
    *ynorm *= lambda;
    return ok ? 0 : 1;
}

typedef struct {
    int its, reason, last_kspits, ksp_total;
    double fnorm0, fnorm;
} obs_stage;

/* SNESSolve_VINEWTONRSLS on grid g from X (in/out); F receives F(X) */
static int rsls(const obs_cfg *cfg, const ogrid *g, double *X, double *F, const double *xl, const double *xu,
                int ind, obs_stage *st, char *out, int outlen, int *pos, char *err, int errlen) {
    const long n = g->n;
    memset(st, 0, sizeof *st);
    oprob pr;
    memset(&pr, 0, sizeof pr);
    pr.dim = 2; pr.cx = pr.cy = pr.cz = 1.0;
    ocsr J;
    oracle_jacobian(g, &pr, &J);
    double *Y = (double *)calloc(n, sizeof(double));
    double *W = (double *)malloc(sizeof(double) * n), *G = (double *)malloc(sizeof(double) * n);
    long *keep = (long *)malloc(sizeof(long) * n), *map = (long *)malloc(sizeof(long) * n);
    double *Fi = (double *)malloc(sizeof(double) * n), *Yi = (double *)malloc(sizeof(double) * n);
    lsctx lc = {g, &J, xl, xu, W, G};
    int e = 0;
    if (cfg->sym) oprintf(out, outlen, pos, "%*sMatrix is symmetric (tolerance %s)\n", ind, "", cfg->symtol);
    vi_project(n, X, xl, xu);
    obs_residual(g, X, F);
    double fnorm = vi_fnorm(n, F, X, xl, xu), xnorm = ovnorm2(n, X), ynorm = 0.0;
    double ttol = 0.0, rnorm0 = 0.0;
    st->fnorm0 = st->fnorm = fnorm;
    int reason = snes_converged(cfg, 0, 0.0, 0.0, fnorm, &ttol, &rnorm0);
    if (cfg->monitor) snes_line(out, outlen, pos, ind, 0, fnorm);
    int i = 0, iter = 0;   /* iter: snes->iter */
    for (; !reason && i < cfg->maxit; i++) {
        if (cfg->sym) oprintf(out, outlen, pos, "%*sMatrix is symmetric (tolerance %s)\n", ind, "", cfg->symtol);
        long nk = 0;
        for (long q = 0; q < n; q++) {
            if (vi_active(X[q], F[q], xl[q], xu ? xu[q] : INFINITY)) map[q] = -1;
            else { map[q] = nk; keep[nk++] = q; }
        }
        ocsr Jii;
        csr_submatrix(&J, nk, keep, map, &Jii);
        for (long r = 0; r < nk; r++) Fi[r] = F[keep[r]];
        int kreason = 3, kits = 0;
        if (nk > 0) {
            vimg mg;
            oasm as;
            int pct = cfg->pctype;
            if (pct == PC_MG) {
                int nlev = cfg->mg_levels;
                e = vimg_setup(&mg, g, nlev, nk, keep, &Jii, cfg->o, err, errlen);
                if (e) { ocsr_free(&Jii); break; }
                pct = PC_SHELL;
            } else if (cfg->pcasm) {
                e = oasm_setup(&as, g, &Jii, nk, keep, cfg->nranks, cfg->asm_overlap, err, errlen);
                if (e) { oasm_free(&as); ocsr_free(&Jii); break; }
            }
            oksp k;
            oksp_init(&k, cfg->ktype, &Jii, pct);
            if (pct == PC_SHELL && cfg->pcasm) { k.pc.shell = oasm_apply; k.pc.shell_ctx = &as; }
            else if (pct == PC_SHELL) { k.pc.shell = vimg_apply; k.pc.shell_ctx = &mg; }
            k.rtol = cfg->krtol; k.atol = cfg->katol; k.dtol = cfg->kdtol; k.maxit = cfg->kmaxit;
            if (cfg->knormset) k.normtype = cfg->normtype;
            double r0, rn;
            kits = oksp_solve(&k, Fi, Yi, &kreason, &r0, &rn);
            oksp_free(&k);
            if (cfg->pctype == PC_MG) vimg_free(&mg);
            if (cfg->pcasm) oasm_free(&as);
        }
        ocsr_free(&Jii);
        ovset(n, 0.0, Y);
        for (long r = 0; r < nk; r++) Y[keep[r]] = Yi[r];
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
        if (kreason < 0) { reason = -3; break; }   /* maxLinearSolveFailures = 1 */
        double gnorm = fnorm;
        const int lsfail = line_search(cfg, &lc, X, F, &gnorm, Y, &xnorm, &ynorm);
        if (lsfail) {   /* maxFailures = 1 */
            reason = (cfg->stol * xnorm > ynorm) ? 4 : -6;
            break;
        }
        fnorm = gnorm;
        iter = i + 1;
        xnorm = ovnorm2(n, X);
        reason = snes_converged(cfg, iter, xnorm, ynorm, fnorm, &ttol, &rnorm0);
        if (cfg->monitor) snes_line(out, outlen, pos, ind, iter, fnorm);
    }
    if (!reason && !e && i >= cfg->maxit) reason = -5;
    st->its = iter;
    st->reason = reason;
    st->fnorm = fnorm;
    ocsr_free(&J);
    free(Y); free(W); free(G); free(keep); free(map); free(Fi); free(Yi);
    return e;
}

/* SNESSolve_VINEWTONSSLS on grid g from X (in/out); F receives F(X).  The
 * semismooth reformulation phi(X) = Phi(X - xl, F(X)) (Fischer-Burmeister);
 * per iteration the Jacobian J, the B-subdifferential vectors Da_i =
 * Phi_a(x_i - l_i, F_i), Db_i = Phi_b = Phi_a(F_i, x_i - l_i), the
 * semismooth Jacobian diag(Db) J + diag(Da) (MatDiagonalScale, then
 * MatDiagonalSet ADD_VALUES), KSPSolve(Jss, phi) for the direction, the
 * line search on |phi|, SNESConvergedDefault on |phi|. */
static int ssls(const obs_cfg *cfg, const ogrid *g, double *X, double *F, const double *xl, int ind,
                obs_stage *st, char *out, int outlen, int *pos) {
    const long n = g->n;
    memset(st, 0, sizeof *st);
    oprob pr;
    memset(&pr, 0, sizeof pr);
    pr.dim = 2; pr.cx = pr.cy = pr.cz = 1.0;
    double *phi = (double *)malloc(sizeof(double) * n), *Y = (double *)calloc(n, sizeof(double));
    double *W = (double *)malloc(sizeof(double) * n), *G = (double *)malloc(sizeof(double) * n);
    double *FW = (double *)malloc(sizeof(double) * n);
    ssls_function(g, X, xl, F, phi);
    double phinorm = ovnorm2(n, phi), xnorm = ovnorm2(n, X), ynorm = 0.0, ttol = 0.0, rnorm0 = 0.0;
    st->fnorm0 = st->fnorm = phinorm;
    int reason = snes_converged(cfg, 0, 0.0, 0.0, phinorm, &ttol, &rnorm0);
    if (cfg->monitor) snes_line(out, outlen, pos, ind, 0, phinorm);
    int i = 0, iter = 0;
    for (; !reason && i < cfg->maxit; i++) {
        if (cfg->sym) oprintf(out, outlen, pos, "%*sMatrix is symmetric (tolerance %s)\n", ind, "", cfg->symtol);
        ocsr J;
        oracle_jacobian(g, &pr, &J);
        for (long r = 0; r < n; r++) {
            const double a = X[r] - xl[r], da = fb_dphi(a, F[r]), db = fb_dphi(F[r], a);
            for (int64_t q = J.rp[r]; q < J.rp[r + 1]; q++) J.v[q] *= db;
            for (int64_t q = J.rp[r]; q < J.rp[r + 1]; q++)
                if (J.ci[q] == r) J.v[q] += da;
        }
        oksp k;
        oksp_init(&k, cfg->ktype, &J, cfg->pctype);
        k.rtol = cfg->krtol; k.atol = cfg->katol; k.dtol = cfg->kdtol; k.maxit = cfg->kmaxit;
        if (cfg->knormset) k.normtype = cfg->normtype;
        int kreason = 0;
        double r0, rn;
        const int kits = oksp_solve(&k, phi, Y, &kreason, &r0, &rn);
        oksp_free(&k);
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
        lsctx lc = {g, &J, xl, NULL, W, G, 1, cfg->ls_project, FW};
        double gnorm = phinorm;
        const int lsfail = line_search(cfg, &lc, X, phi, &gnorm, Y, &xnorm, &ynorm);
        ocsr_free(&J);
        memcpy(F, FW, sizeof(double) * n);   /* the residual at the accepted point */
        if (lsfail) {
            reason = (cfg->stol * xnorm > ynorm) ? 4 : -6;
            break;
        }
        phinorm = gnorm;
        iter = i + 1;
        xnorm = ovnorm2(n, X);
        reason = snes_converged(cfg, iter, xnorm, ynorm, phinorm, &ttol, &rnorm0);
        if (cfg->monitor) snes_line(out, outlen, pos, ind, iter, phinorm);
    }
    if (!reason && i >= cfg->maxit) reason = -5;
    st->its = iter;
    st->reason = reason;
    st->fnorm = phinorm;
    free(phi); free(Y); free(W); free(G); free(FW);
    return 0;
}

/* --------------------------------------------------------------- driver */

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int oracle_obstacle_run(int argc, const char **argv, obstacle_result *res, double *u_out, char *out, int outlen) {
    obstacle_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    if (out && outlen > 0) out[0] = 0;
    int pos = 0;
    char err[512] = "";
    oopts o;
    oopts_parse(&o, argc, argv);
    obs_cfg c;
    memset(&c, 0, sizeof c);
    c.o = &o;
    const char *st = oopts_get(&o, "snes_type");
    if (st && strcmp(st, "vinewtonrsls") && strcmp(st, "vinewtonssls")) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: -snes_type %s not supported\n", st);
        return 56;
    }
    c.ssls = st && !strcmp(st, "vinewtonssls");
    c.ls_alpha = oopts_real(&o, "oracle_ls_alpha", 0.0);
    c.ls_project = oopts_int(&o, "oracle_ls_project", 1);   /* 0: a variant the tests show test4 rules out */
    const char *lt = oopts_get(&o, "snes_linesearch_type");
    c.ls = (lt && !strcmp(lt, "basic")) ? LS_BASIC : LS_BT;
    if (lt && strcmp(lt, "basic") && strcmp(lt, "bt")) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: -snes_linesearch_type %s not supported\n", lt);
        return 56;
    }
    c.ktype = oracle_parse_ksp(oopts_get(&o, "ksp_type"), KSP_CG);   /* obstacle.c:126 */
    const char *pcs = oopts_get(&o, "pc_type");
    c.pcasm = pcs && !strcmp(pcs, "asm");
    c.pctype = c.pcasm ? PC_SHELL : oracle_parse_pc(pcs, PC_ILU);
    c.nranks = oopts_int(&o, "oracle_nranks", 1);
    c.asm_overlap = oopts_int(&o, "pc_asm_overlap", 1);
    if (c.pcasm) {
        const char *sp = oopts_get(&o, "sub_pc_type");
        if (!sp || strcmp(sp, "lu") || c.nranks < 1 || c.nranks > 16 || c.asm_overlap < 0) {
            oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: -pc_type asm needs -sub_pc_type lu, 1..16 ranks\n");
            return 56;
        }
    }
    if (c.ktype < 0 || c.pctype < 0 || c.ktype == KSP_PREONLY) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: oracle: unsupported -ksp_type/-pc_type\n");
        return 56;
    }
    c.krtol = oopts_real(&o, "ksp_rtol", 1e-5);
    c.katol = oopts_real(&o, "ksp_atol", 1e-50);
    c.kdtol = oopts_real(&o, "ksp_divtol", 1e5);
    c.kmaxit = oopts_int(&o, "ksp_max_it", 10000);
    const char *nt = oopts_get(&o, "ksp_norm_type");
    if (nt) {
        c.knormset = 1;
        if (!strcmp(nt, "preconditioned")) c.normtype = NORM_PRECONDITIONED;
        else if (!strcmp(nt, "unpreconditioned")) c.normtype = NORM_UNPRECONDITIONED;
        else if (!strcmp(nt, "natural")) c.normtype = NORM_NATURAL;
        else if (!strcmp(nt, "none")) c.normtype = NORM_NONE;
        else c.knormset = 0;
    }
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
    /* grid (obstacle.c:86-95): 3x3 base, -da_grid_x/y, -da_refine */
    int mx = oopts_int(&o, "da_grid_x", 3), my = oopts_int(&o, "da_grid_y", 3);
    const int refine = oopts_int(&o, "da_refine", 0), seq = oopts_int(&o, "snes_grid_sequence", 0);
    if (seq < 0 || seq > 12) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: -snes_grid_sequence %d out of range\n", seq);
        return 63;
    }
    for (int r = 0; r < refine; r++) { mx = 2 * (mx - 1) + 1; my = 2 * (my - 1) + 1; }
    const int lev_opt = oopts_int(&o, "pc_mg_levels", -1);
    ogrid g;
    ogrid_set(&g, 2, mx, my, 1, 4.0, 4.0, 1.0);
    double *X = (double *)calloc(g.n, sizeof(double));   /* obstacle.c:130-131 */
    double t0 = now_ms();
    obs_stage sg;
    int e = 0;
    res->stages = seq + 1;
    for (int stage = 0; stage <= seq; stage++) {
        const long n = g.n;
        double *xl = (double *)malloc(sizeof(double) * n), *F = (double *)malloc(sizeof(double) * n);
        obs_bounds(&g, xl);
        c.mg_levels = lev_opt > 0 ? lev_opt : refine + stage + 1;
        const int ind = 2 * (seq - stage);
        if (c.ssls) {
            if (c.pctype == PC_MG) {
                snprintf(err, sizeof err, "oracle: -snes_type vinewtonssls with -pc_type mg");
                e = 56;
            } else {
                e = ssls(&c, &g, X, F, xl, ind, &sg, out, outlen, &pos);
            }
        } else {
            e = rsls(&c, &g, X, F, xl, NULL, ind, &sg, out, outlen, &pos, err, sizeof err);
        }
        if (e) { free(xl); free(F); break; }
        if (stage < 16) res->stage_its[stage] = sg.its;
        if (snes_reason) {
            if (sg.reason > 0)
                oprintf(out, outlen, &pos, "%*s  Nonlinear solve converged due to %s iterations %d\n", ind, "",
                        snes_reason_name(sg.reason), sg.its);
            else
                oprintf(out, outlen, &pos, "%*s  Nonlinear solve did not converge due to %s iterations %d\n", ind,
                        "", snes_reason_name(sg.reason), sg.its);
        }
        if (stage < seq) {   /* DMRefine + MatInterpolate */
            ogrid gf;
            ogrid_set(&gf, 2, 2 * (g.mx - 1) + 1, 2 * (g.my - 1) + 1, 1, 4.0, 4.0, 1.0);
            double *Xf = (double *)calloc(gf.n, sizeof(double));
            oq1_interp_add(&g, &gf, X, Xf);
            free(X);
            X = Xf;
            g = gf;
            free(xl); free(F);
            continue;
        }
        /* final report (obstacle.c:160-198) */
        res->mx = g.mx; res->my = g.my;
        res->its = sg.its; res->reason = sg.reason; res->last_ksp_its = sg.last_kspits;
        res->ksp_its_total = sg.ksp_total; res->fnorm0 = sg.fnorm0; res->fnorm = sg.fnorm;
        if (sg.reason <= 0)
            oprintf(out, outlen, &pos, "WARNING: SNES not converged ... use -snes_converged_reason to check\n");
        oprintf(out, outlen, &pos, "done on %d x %d grid ... %s, SNES iters = %d, last KSP iters = %d\n", g.mx,
                g.my, snes_reason_name(sg.reason), sg.its, sg.last_kspits);
        long act = 0;
        double e1 = 0.0, einf = 0.0;
        for (int j = 0; j < g.my; j++)
            for (int i = 0; i < g.mx; i++) {
                const long q = (long)j * g.mx + i;
                if (X[q] <= xl[q] + VI_ZEROTOL && F[q] > 0.0) act++;
                const double d = fabs(X[q] - obs_uexact(-2.0 + i * g.hx, -2.0 + j * g.hy));
                e1 += d;
                if (d > einf) einf = d;
            }
        const double dx = 4.0 / (g.mx - 1), dy = 4.0 / (g.my - 1);
        const double actarea = dx * dy * act, exactarea = M_PI * OBS_AEXACT * OBS_AEXACT;
        res->active = act;
        res->error1 = e1 / ((double)g.mx * g.my);
        res->errorinf = einf;
        res->areaerr = fabs(actarea - exactarea) / exactarea;
        oprintf(out, outlen, &pos, "errors: av |u-uexact| = %.3e, |u-uexact|_inf = %.3e, active area error = %.3f%%\n",
                res->error1, res->errorinf, 100.0 * res->areaerr);
        if (u_out) memcpy(u_out, X, sizeof(double) * n);
        free(xl); free(F);
    }
    res->solve_ms = now_ms() - t0;
    if (e) oprintf(out, outlen, &pos, "[0]PETSC ERROR: %s\n", err);
    free(X);
    return e;
}
