/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * Krylov methods, preconditioners and the multigrid V/W-cycle, restating
 * the documented semantics of the PETSc objects the reference path selects
 * with `-ksp_type cg -pc_type mg` (c/ch6/fish.c:231-234, c/ch6/makefile:12-33):
 *   KSPCG (preconditioned-norm PCG, default rtol 1e-5, atol 1e-50, dtol 1e5),
 *   PCMG multiplicative cycles with one KSP per level used for pre- and
 *   post-smoothing (nonzero initial guess), x_c = 0 before recursion,
 *   restriction = P^T, interpolation-add = P; PCMG type full (PETSc's
 *   PCMGFCycle_Private: the right-hand side restricted to the coarsest
 *   level, then upward a cycle per level started from the interpolated
 *   coarser iterate, x_{l+1} = P x_l),
 *   KSPCHEBYSHEV (first step x += s B r, then the three-term recurrence;
 *   bounds [0.1, 1.1] x estimated lambda_max from 10 Krylov steps),
 *   KSPRICHARDSON, PCJACOBI, PCSOR (symmetric local sweeps, omega 1),
 *   PCILU(0) natural ordering, PCLU coarse solve.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

const char *oracle_reason_name(int r) {
    switch (r) {
    case R_CONVERGED_RTOL: return "CONVERGED_RTOL";
    case R_CONVERGED_ATOL: return "CONVERGED_ATOL";
    case R_CONVERGED_ITS: return "CONVERGED_ITS";
    case R_DIVERGED_NULL: return "DIVERGED_NULL";
    case R_DIVERGED_ITS: return "DIVERGED_ITS";
    case R_DIVERGED_DTOL: return "DIVERGED_DTOL";
    case R_DIVERGED_BREAKDOWN: return "DIVERGED_BREAKDOWN";
    case R_DIVERGED_INDEFINITE_PC: return "DIVERGED_INDEFINITE_PC";
    case R_DIVERGED_INDEFINITE_MAT: return "DIVERGED_INDEFINITE_MAT";
    case R_DIVERGED_NANORINF: return "DIVERGED_NANORINF";
    }
    return "CONVERGED_ITERATING";
}

/* ------------------------------------------------------------- PC setup */

static int64_t diag_pos(const ocsr *A, long i) {
    for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++)
        if (A->ci[q] == i) return q;
    return -1;
}

static void ilu0_factor(opc *pc) {
    const ocsr *A = pc->A;
    long n = A->n;
    int64_t nnz = A->rp[n];
    pc->F.n = n;
    pc->F.rp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    pc->F.ci = (int32_t *)malloc(sizeof(int32_t) * nnz);
    pc->F.v = (double *)malloc(sizeof(double) * nnz);
    memcpy(pc->F.rp, A->rp, sizeof(int64_t) * (n + 1));
    memcpy(pc->F.ci, A->ci, sizeof(int32_t) * nnz);
    memcpy(pc->F.v, A->v, sizeof(double) * nnz);
    pc->dpos = (int64_t *)malloc(sizeof(int64_t) * n);
    for (long i = 0; i < n; i++) pc->dpos[i] = diag_pos(&pc->F, i);
    ocsr *F = &pc->F;
    /* IKJ ILU(0): same sparsity as A */
    for (long i = 0; i < n; i++) {
        for (int64_t q = F->rp[i]; q < pc->dpos[i]; q++) {
            long k = F->ci[q];
            double l = F->v[q] / F->v[pc->dpos[k]];
            F->v[q] = l;
            for (int64_t q2 = pc->dpos[k] + 1; q2 < F->rp[k + 1]; q2++) {
                long j = F->ci[q2];
                for (int64_t q3 = q + 1; q3 < F->rp[i + 1]; q3++)
                    if (F->ci[q3] == j) { F->v[q3] -= l * F->v[q2]; break; }
            }
        }
    }
}

static void band_factor(opc *pc) {
    const ocsr *A = pc->A;
    long n = A->n;
    int bw = 0;
    for (long i = 0; i < n; i++)
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
            long d = labs((long)A->ci[q] - i);
            if (d > bw) bw = (int)d;
        }
    pc->bw = bw;
    long w = 2 * bw + 1;
    pc->band = (double *)calloc((size_t)n * w, sizeof(double));
#define B(i, j) pc->band[(long)(i) * w + ((j) - (i) + bw)]
    for (long i = 0; i < n; i++)
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) B(i, A->ci[q]) = A->v[q];
    /* LU without pivoting (A is SPD) */
    for (long k = 0; k < n; k++) {
        double piv = B(k, k);
        long iend = (k + bw < n - 1) ? k + bw : n - 1;
        for (long i = k + 1; i <= iend; i++) {
            double l = B(i, k) / piv;
            if (l == 0.0) continue;
            B(i, k) = l;
            for (long j = k + 1; j <= iend; j++) B(i, j) -= l * B(k, j);
        }
    }
#undef B
}

static void band_solve(const opc *pc, const double *r, double *z) {
    long n = pc->A->n, w = 2 * pc->bw + 1;
    int bw = pc->bw;
    const double *band = pc->band;
#define B(i, j) band[(long)(i) * w + ((j) - (i) + bw)]
    for (long i = 0; i < n; i++) {
        double s = r[i];
        long j0 = (i - bw > 0) ? i - bw : 0;
        for (long j = j0; j < i; j++) s -= B(i, j) * z[j];
        z[i] = s;
    }
    for (long i = n - 1; i >= 0; i--) {
        double s = z[i];
        long j1 = (i + bw < n - 1) ? i + bw : n - 1;
        for (long j = i + 1; j <= j1; j++) s -= B(i, j) * z[j];
        z[i] = s / B(i, i);
    }
#undef B
}

void opc_setup(opc *pc, int type, const ocsr *A) {
    memset(pc, 0, sizeof *pc);
    pc->type = type;
    pc->A = A;
    pc->omega = 1.0;
    pc->nblk = 1; pc->blk[0] = 0; pc->blk[1] = A->n;
    if (type == PC_JACOBI || type == PC_SOR) {
        pc->idiag = (double *)malloc(sizeof(double) * A->n);
        for (long i = 0; i < A->n; i++) {
            int64_t q = diag_pos(A, i);
            pc->idiag[i] = 1.0 / A->v[q];
        }
    } else if (type == PC_ILU) {
        ilu0_factor(pc);
    } else if (type == PC_LU) {
        band_factor(pc);
    }
}

void opc_free(opc *pc) {
    free(pc->idiag);
    if (pc->F.v) ocsr_free(&pc->F);
    free(pc->dpos);
    free(pc->band);
    if (pc->mg) { omg_free(pc->mg); free(pc->mg); }
    memset(pc, 0, sizeof *pc);
}

/* MatSOR, zero initial guess, local symmetric sweep, omega 1:
 * forward x = (D+L)^-1 b, x <- D x, backward x = (D+U)^-1 x (per block) */
static void sor_zero_guess(const opc *pc, const double *b, double *x) {
    const ocsr *A = pc->A;
    for (int bk = 0; bk < pc->nblk; bk++) {
        long lo = pc->blk[bk], hi = pc->blk[bk + 1];
        for (long i = lo; i < hi; i++) {
            double s = b[i];
            for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
                long j = A->ci[q];
                if (j >= i) break;
                if (j >= lo) s -= A->v[q] * x[j];
            }
            x[i] = s * pc->idiag[i];
        }
        for (long i = lo; i < hi; i++) x[i] *= 1.0 / pc->idiag[i];
        for (long i = hi - 1; i >= lo; i--) {
            double s = x[i];
            for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
                long j = A->ci[q];
                if (j <= i || j >= hi) continue;
                s -= A->v[q] * x[j];
            }
            x[i] = s * pc->idiag[i];
        }
    }
}

/* symmetric sweeps from a nonzero guess: x_i += (b_i - (A x)_i)/a_ii,
 * off-block couplings frozen at their values on entry (local sweeps)    */
void opc_sor_richardson(opc *pc, const double *b, double *x, int its) {
    const ocsr *A = pc->A;
    long n = A->n;
    double *bb = (double *)malloc(sizeof(double) * n);
    for (int it = 0; it < its; it++) {
        for (int bk = 0; bk < pc->nblk; bk++) {
            long lo = pc->blk[bk], hi = pc->blk[bk + 1];
            for (long i = lo; i < hi; i++) {
                double s = b[i];
                for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
                    long j = A->ci[q];
                    if (j < lo || j >= hi) s -= A->v[q] * x[j];
                }
                bb[i] = s;
            }
        }
        for (int bk = 0; bk < pc->nblk; bk++) {
            long lo = pc->blk[bk], hi = pc->blk[bk + 1];
            for (long i = lo; i < hi; i++) {
                double s = bb[i], d = 0.0;
                for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
                    long j = A->ci[q];
                    if (j < lo || j >= hi) continue;
                    if (j == i) d = A->v[q];
                    s -= A->v[q] * x[j];
                }
                x[i] = (1.0 - pc->omega) * x[i] + (s + d * x[i]) * (pc->omega / d);
            }
            for (long i = hi - 1; i >= lo; i--) {
                double s = bb[i], d = 0.0;
                for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) {
                    long j = A->ci[q];
                    if (j < lo || j >= hi) continue;
                    if (j == i) d = A->v[q];
                    s -= A->v[q] * x[j];
                }
                x[i] = (1.0 - pc->omega) * x[i] + (s + d * x[i]) * (pc->omega / d);
            }
        }
    }
    free(bb);
}

void opc_apply(opc *pc, const double *r, double *z) {
    long n = pc->A->n;
    switch (pc->type) {
    case PC_NONE: ovcopy(n, r, z); break;
    case PC_JACOBI:
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) z[i] = r[i] * pc->idiag[i];
        break;
    case PC_SOR: sor_zero_guess(pc, r, z); break;
    case PC_ILU: {
        const ocsr *F = &pc->F;
        for (long i = 0; i < n; i++) {
            double s = r[i];
            for (int64_t q = F->rp[i]; q < pc->dpos[i]; q++) s -= F->v[q] * z[F->ci[q]];
            z[i] = s;
        }
        for (long i = n - 1; i >= 0; i--) {
            double s = z[i];
            for (int64_t q = pc->dpos[i] + 1; q < F->rp[i + 1]; q++) s -= F->v[q] * z[F->ci[q]];
            z[i] = s / F->v[pc->dpos[i]];
        }
        break;
    }
    case PC_LU: band_solve(pc, r, z); break;
    case PC_MG: omg_apply(pc->mg, r, z); break;
    case PC_SHELL: pc->shell(pc->shell_ctx, r, z); break;
    }
}

/* ------------------------------------------------------------------ KSP */

void oksp_init(oksp *k, int type, const ocsr *A, int pctype) {
    memset(k, 0, sizeof *k);
    k->type = type;
    k->A = A;
    k->pctype = pctype;
    k->rtol = 1e-5; k->atol = 1e-50; k->dtol = 1e5;
    k->maxit = 10000;
    k->rich_scale = 1.0;
    k->normtype = (type == KSP_CG || type == KSP_GMRES) ? NORM_PRECONDITIONED : NORM_NONE;
    k->restart = 30;
    k->esteig_steps = 10;
    k->est_tf[0] = 0.0; k->est_tf[1] = 0.1; k->est_tf[2] = 0.0; k->est_tf[3] = 1.1;
    k->est_lanczos = 1;
    for (int q = 0; q < 5; q++) k->w[q] = (double *)malloc(sizeof(double) * A->n);
    if (pctype != PC_MG) opc_setup(&k->pc, pctype, A);
    else { k->pc.type = PC_MG; k->pc.A = A; }
}

/* the operator of a KSP: its matrix, or a shell product (amult) such as a
 * matrix-free finite-difference Jacobian (MatMFFD) */
static void kmult(const oksp *k, const double *x, double *y) {
    if (k->amult) k->amult(k->actx, x, y);
    else ocsr_mult(k->A, x, y);
}

/* r = b - A x (KSPInitialResidual: the product, then VecAYPX) */
static void kresid(const oksp *k, const double *b, const double *x, double *r) {
    if (!k->amult) { ocsr_resid(k->A, b, x, r); return; }
    k->amult(k->actx, x, r);
    for (long i = 0; i < k->A->n; i++) r[i] = b[i] - r[i];
}

void oksp_free(oksp *k) {
    for (int q = 0; q < 5; q++) free(k->w[q]);
    opc_free(&k->pc);
}

/* symmetric tridiagonal (diag a[0..m-1], offdiag b[0..m-2]) extreme
 * eigenvalues by the cyclic Jacobi method on the dense m x m matrix */
static void tridiag_extreme(int m, const double *a, const double *b, double *emin, double *emax) {
    double T[16][16];
    memset(T, 0, sizeof T);
    for (int i = 0; i < m; i++) {
        T[i][i] = a[i];
        if (i + 1 < m) T[i][i + 1] = T[i + 1][i] = b[i];
    }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) off += T[p][q] * T[p][q];
        if (off < 1e-30) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                if (fabs(T[p][q]) < 1e-300) continue;
                double th = (T[q][q] - T[p][p]) / (2.0 * T[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < m; r++) {
                    double trp = T[r][p], trq = T[r][q];
                    T[r][p] = c * trp - s * trq;
                    T[r][q] = s * trp + c * trq;
                }
                for (int r = 0; r < m; r++) {
                    double tpr = T[p][r], tqr = T[q][r];
                    T[p][r] = c * tpr - s * tqr;
                    T[q][r] = s * tpr + c * tqr;
                }
            }
    }
    *emin = T[0][0]; *emax = T[0][0];
    for (int i = 1; i < m; i++) {
        if (T[i][i] < *emin) *emin = T[i][i];
        if (T[i][i] > *emax) *emax = T[i][i];
    }
}

/* Chebyshev eigen-estimate: `esteig_steps` steps of PCG (zero guess) on the
 * deterministic noisy right-hand side; the CG coefficients give the
 * Lanczos matrix T_k, whose extreme eigenvalues are mapped by est_tf.     */
void oksp_estimate_eigs(oksp *k) {
    const ocsr *A = k->A;
    long n = A->n;
    double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) r[i] = oracle_noisy(k->gidx ? k->gidx[i] : i);
    double al[16], be[16];
    int m = 0;
    opc_apply(&k->pc, r, z);
    double rz = ovdot(n, r, z), rzold = 0.0;
    ovcopy(n, z, p);
    for (int it = 0; it < k->esteig_steps; it++) {
        if (!(rz > 0.0)) break;
        if (it > 0) {
            double beta = rz / rzold;
            be[it - 1] = beta;
            for (long i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        }
        kmult(k, p, w);
        double pw = ovdot(n, p, w);
        if (!(pw > 0.0)) break;
        double alpha = rz / pw;
        al[it] = alpha;
        m = it + 1;
        for (long i = 0; i < n; i++) r[i] -= alpha * w[i];
        opc_apply(&k->pc, r, z);
        rzold = rz;
        rz = ovdot(n, r, z);
        /* stop on (numerical) exhaustion of the Krylov space */
        if (rz <= 1e-28 * fabs(rzold)) break;
    }
    double a[16], b[16], emin = 1.0, emax = 1.0;
    for (int j = 0; j < m; j++) {
        a[j] = 1.0 / al[j] + (j > 0 ? be[j - 1] / al[j - 1] : 0.0);
        if (j + 1 < m) b[j] = sqrt(be[j]) / al[j];
    }
    if (m > 0) tridiag_extreme(m, a, b, &emin, &emax);
    k->emin = k->est_tf[0] * emin + k->est_tf[1] * emax;
    k->emax = k->est_tf[2] * emin + k->est_tf[3] * emax;
    k->est_done = 1;
    free(r); free(z); free(p); free(w);
}

static void ksp_monitor(oksp *k, int it, double rn) {
    if (!k->monitor) return;
    oprintf(k->mbuf, k->mlen, k->mpos, "%s%3d KSP Residual norm %14.12e \n",
            k->prefix ? k->prefix : "    ", it, rn);
}

static int converged(oksp *k, int it, double rn, double *rnorm0, double *ttol) {
    if (isnan(rn) || isinf(rn)) return R_DIVERGED_NANORINF;
    if (it == 0) {
        *rnorm0 = rn;
        *ttol = k->rtol * rn > k->atol ? k->rtol * rn : k->atol;
    }
    if (rn <= *ttol) return (rn < k->atol) ? R_CONVERGED_ATOL : R_CONVERGED_RTOL;
    if (it > 0 && rn >= k->dtol * *rnorm0) return R_DIVERGED_DTOL;
    return 0;
}

/* KSPSolve_CG (preconditioned norm by default) */
static int cg_solve(oksp *k, const double *b, double *x, int *reason, double *rnorm0, double *rnorm) {
    const ocsr *A = k->A;
    long n = A->n;
    double *r = k->w[0], *z = k->w[1], *p = k->w[2], *w = k->w[3];
    double dp = 0.0, beta = 0.0, betaold = 0.0, dpi = 0.0, dpiold = 0.0, ttol = 0.0;
    if (k->guess_nonzero) kresid(k, b, x, r);
    else { ovset(n, 0.0, x); ovcopy(n, b, r); }
    if (k->normtype == NORM_PRECONDITIONED) { opc_apply(&k->pc, r, z); dp = ovnorm2(n, z); }
    else if (k->normtype == NORM_UNPRECONDITIONED) dp = ovnorm2(n, r);
    else if (k->normtype == NORM_NATURAL) { opc_apply(&k->pc, r, z); beta = ovdot(n, z, r); dp = sqrt(fabs(beta)); }
    ksp_monitor(k, 0, dp);
    *reason = (k->normtype == NORM_NONE) ? 0 : converged(k, 0, dp, rnorm0, &ttol);
    *rnorm = dp;
    if (*reason) return 0;
    if (k->normtype != NORM_PRECONDITIONED && k->normtype != NORM_NATURAL) opc_apply(&k->pc, r, z);
    if (k->normtype != NORM_NATURAL) beta = ovdot(n, z, r);
    int i = 0, its = 0;
    do {
        its = i + 1;
        if (beta == 0.0) { *reason = R_CONVERGED_ATOL; break; }
        if (i > 0 && beta * betaold < 0.0) { *reason = R_DIVERGED_INDEFINITE_PC; break; }
        if (i == 0) ovcopy(n, z, p);
        else {
            double bb = beta / betaold;
#pragma omp parallel for schedule(static)
            for (long q = 0; q < n; q++) p[q] = z[q] + bb * p[q];
        }
        dpiold = dpi;
        kmult(k, p, w);
        dpi = ovdot(n, p, w);
        betaold = beta;
        if (dpi == 0.0 || (i > 0 && ((dpi > 0) != (dpiold > 0)))) {
            *reason = R_DIVERGED_INDEFINITE_MAT; break;
        }
        double a = beta / dpi;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; q++) { x[q] += a * p[q]; r[q] -= a * w[q]; }
        if (k->normtype == NORM_PRECONDITIONED) { opc_apply(&k->pc, r, z); dp = ovnorm2(n, z); }
        else if (k->normtype == NORM_UNPRECONDITIONED) dp = ovnorm2(n, r);
        else if (k->normtype == NORM_NATURAL) { opc_apply(&k->pc, r, z); beta = ovdot(n, z, r); dp = sqrt(fabs(beta)); }
        *rnorm = dp;
        ksp_monitor(k, i + 1, dp);
        if (k->normtype != NORM_NONE) *reason = converged(k, i + 1, dp, rnorm0, &ttol);
        if (*reason) break;
        if (k->normtype != NORM_PRECONDITIONED && k->normtype != NORM_NATURAL) opc_apply(&k->pc, r, z);
        if (k->normtype != NORM_NATURAL) beta = ovdot(n, z, r);
        i++;
    } while (i < k->maxit);
    if (i >= k->maxit && !*reason) *reason = R_DIVERGED_ITS;
    return its;
}

/* KSPSolve_GMRES (PETSc's default KSP for SNES): restarted GMRES(m), left
 * preconditioning, classical Gram-Schmidt without refinement, the residual
 * norm of the preconditioned system from the Givens-rotated Hessenberg
 * least-squares problem.  A cycle starts from the true residual B (b - A x);
 * the convergence test runs at the start of each cycle (ksp its so far) and
 * after each Arnoldi step; the happy breakdown tolerance is 1e-30.  H is the
 * rotated Hessenberg column-major with leading dimension m + 2.            */
static int gmres_solve(oksp *k, const double *b, double *x, int *reason, double *rnorm0, double *rnorm) {
    const ocsr *A = k->A;
    const long n = A->n;
    const int m = k->restart > 0 ? k->restart : 30;
    const int ld = m + 2;
    double **V = (double **)malloc(sizeof(double *) * (m + 1));
    for (int j = 0; j <= m; j++) V[j] = (double *)malloc(sizeof(double) * (n ? n : 1));
    double *H = (double *)calloc((size_t)ld * (m + 1), sizeof(double));
    double *cc = (double *)calloc(m + 1, sizeof(double)), *ss = (double *)calloc(m + 1, sizeof(double));
    double *grs = (double *)calloc(m + 2, sizeof(double)), *h = (double *)malloc(sizeof(double) * (m + 1));
    double *t = k->w[0], *aw = k->w[1], *corr = k->w[2];
    double ttol = 0.0, rn = -1.0;
    int its = 0, first = 1;
    *reason = 0;
#define HH(a, c) H[(size_t)(c) * ld + (a)]
    while (!*reason) {
        /* KSPInitialResidual: V0 = B (b - A x), or B b from a zero guess */
        if (first && !k->guess_nonzero) { ovset(n, 0.0, x); opc_apply(&k->pc, b, V[0]); }
        else { kresid(k, b, x, t); opc_apply(&k->pc, t, V[0]); }
        /* KSPGMRESCycle */
        double res = ovnorm2(n, V[0]);
        if (res != 0.0) { const double s = 1.0 / res; for (long i = 0; i < n; i++) V[0][i] *= s; }
        grs[0] = res;
        rn = res;
        ksp_monitor(k, its, res);
        if (res == 0.0) { *reason = R_CONVERGED_ATOL; if (first) *rnorm0 = 0.0; break; }
        {
            const int c = converged(k, its, res, rnorm0, &ttol);
            *reason = c;
        }
        first = 0;
        int it = 0, hapend = 0;
        while (!*reason && it < m && its < k->maxit) {
            if (it) ksp_monitor(k, its, res);
            /* V[it+1] = B A V[it] */
            kmult(k, V[it], aw);
            opc_apply(&k->pc, aw, V[it + 1]);
            /* classical Gram-Schmidt: h = V^T v, v -= V h */
            for (int j = 0; j <= it; j++) h[j] = ovdot(n, V[it + 1], V[j]);
            for (long i = 0; i < n; i++) {
                double s = V[it + 1][i];
                for (int j = 0; j <= it; j++) s -= h[j] * V[j][i];
                V[it + 1][i] = s;
            }
            for (int j = 0; j <= it; j++) HH(j, it) = h[j];
            double tt = ovnorm2(n, V[it + 1]);
            if (tt != 0.0) { const double s = 1.0 / tt; for (long i = 0; i < n; i++) V[it + 1][i] *= s; }
            HH(it + 1, it) = tt;
            double hapbnd = fabs(tt / grs[it]);
            if (hapbnd > 1e-30) hapbnd = 1e-30;
            if (tt < hapbnd) hapend = 1;
            /* KSPGMRESUpdateHessenberg: the earlier rotations, then a new one */
            for (int j = 1; j <= it; j++) {
                const double a0 = HH(j - 1, it), a1 = HH(j, it);
                HH(j - 1, it) = cc[j - 1] * a0 + ss[j - 1] * a1;
                HH(j, it) = cc[j - 1] * a1 - ss[j - 1] * a0;
            }
            if (!hapend) {
                const double a0 = HH(it, it), a1 = HH(it + 1, it);
                const double r = sqrt(a0 * a0 + a1 * a1);
                if (r == 0.0) { *reason = R_DIVERGED_NULL; break; }
                cc[it] = a0 / r;
                ss[it] = a1 / r;
                grs[it + 1] = -(ss[it] * grs[it]);
                grs[it] = cc[it] * grs[it];
                HH(it, it) = cc[it] * a0 + ss[it] * a1;
                res = fabs(grs[it + 1]);
            } else {
                res = 0.0;
            }
            it++;
            its++;
            rn = res;
            *reason = converged(k, its, res, rnorm0, &ttol);
            if (hapend && !*reason) { *reason = R_DIVERGED_BREAKDOWN; break; }
        }
        if (it && (*reason || its >= k->maxit)) ksp_monitor(k, its, res);
        /* KSPGMRESBuildSoln: back substitution on the rotated system, x += V y */
        if (it > 0) {
            const int top = it - 1;
            double *y = grs;
            if (HH(top, top) != 0.0) y[top] = grs[top] / HH(top, top);
            else { *reason = R_DIVERGED_BREAKDOWN; y[top] = 0.0; }
            for (int ii = 1; ii <= top; ii++) {
                const int kk = top - ii;
                double s = grs[kk];
                for (int j = kk + 1; j <= top; j++) s -= HH(kk, j) * y[j];
                if (HH(kk, kk) == 0.0) { *reason = R_DIVERGED_BREAKDOWN; break; }
                y[kk] = s / HH(kk, kk);
            }
            for (long i = 0; i < n; i++) {
                double s = 0.0;
                for (int j = 0; j <= top; j++) s += y[j] * V[j][i];
                corr[i] = s;
            }
            for (long i = 0; i < n; i++) x[i] += corr[i];
        }
        if (its >= k->maxit && !*reason) *reason = R_DIVERGED_ITS;
    }
#undef HH
    *rnorm = rn;
    for (int j = 0; j <= m; j++) free(V[j]);
    free(V); free(H); free(cc); free(ss); free(grs); free(h);
    return its;
}

/* KSPSolve_Chebyshev: p[km1] = x, p[k] = x + s B (b - A x), then
 * p[kp1] = (1-w) p[km1] + w p[k] + w s B (b - A p[k])               */
static int cheb_solve(oksp *k, const double *b, double *x) {
    const ocsr *A = k->A;
    long n = A->n;
    if (!k->est_done) oksp_estimate_eigs(k);
    double emax = k->emax, emin = k->emin;
    double scale = 2.0 / (emax + emin);
    double alpha = 1.0 - scale * emin, Gamma = 1.0, mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
    double c[3];
    double *p[3] = {x, k->w[0], k->w[1]}, *r = k->w[2];
    int km1 = 0, kk = 1, kp1 = 2;
    c[km1] = 1.0; c[kk] = mu;
    if (k->guess_nonzero) kresid(k, b, x, r);
    else { ovset(n, 0.0, x); ovcopy(n, b, r); }
    opc_apply(&k->pc, r, p[kk]);
    {
        double *pk = p[kk], *pkm1 = p[km1];
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; q++) pk[q] = scale * pk[q] + pkm1[q];
    }
    for (int i = 1; i < k->maxit; i++) {
        kresid(k, b, p[kk], r);
        opc_apply(&k->pc, r, p[kp1]);
        c[kp1] = 2.0 * mu * c[kk] - c[km1];
        double omega = omegaprod * c[kk] / c[kp1];
        double *a0 = p[km1], *a1 = p[kk], *a2 = p[kp1];
        double ca = 1.0 - omega, cb = omega, cc = omega * Gamma * scale;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; q++) a2[q] = ca * a0[q] + cb * a1[q] + cc * a2[q];
        int t = km1; km1 = kk; kk = kp1; kp1 = t;
    }
    if (kk) ovcopy(n, p[kk], x);
    return k->maxit;
}

/* KSPSolve_Richardson (PCApplyRichardson for SOR) */
static int rich_solve(oksp *k, const double *b, double *x) {
    const ocsr *A = k->A;
    long n = A->n;
    if (k->pctype == PC_SOR && k->rich_scale == 1.0) {
        if (!k->guess_nonzero) ovset(n, 0.0, x);
        opc_sor_richardson(&k->pc, b, x, k->maxit);
        return k->maxit;
    }
    double *r = k->w[0], *z = k->w[1];
    for (int i = 0; i < k->maxit; i++) {
        if (i == 0 && !k->guess_nonzero) { ovset(n, 0.0, x); ovcopy(n, b, r); }
        else kresid(k, b, x, r);
        opc_apply(&k->pc, r, z);
        for (long q = 0; q < n; q++) x[q] += k->rich_scale * z[q];
    }
    return k->maxit;
}

static void mcycle(omg *mg, int l);

/* PCApplyRichardson_MG: up to maxit multiplicative cycles from x = 0, each
 * stopping after the finest pre-smoothing once |b - A x| <= ttol, ttol =
 * max(rtol |b|, atol); returns the number of completed cycles */
static int rich_mg(oksp *k, const double *b, double *x, int *reason) {
    omg *mg = k->pc.mg;
    const int L = mg->nlev - 1;
    const long n = k->A->n;
    double ttol = 0.0;
    if (k->rtol > 0.0) {
        double rn = ovnorm2(n, b);
        ttol = k->rtol * rn > k->atol ? k->rtol * rn : k->atol;
    } else if (k->atol > 0.0) ttol = k->atol;
    int rr = 0;
    mg->ttol = ttol; mg->atol = k->atol; mg->rreason = &rr;
    ovset(n, 0.0, x);
    mg->b[L] = (double *)b; mg->x[L] = x;
    int i = 0;
    for (; i < k->maxit; i++) {
        mcycle(mg, L);
        if (rr) break;
    }
    mg->rreason = NULL; mg->b[L] = NULL; mg->x[L] = NULL;
    *reason = rr ? rr : R_CONVERGED_ITS;
    return i;
}

/* KSPSolve_Richardson, the generic loop from x = 0 (PETSc takes it when a
 * monitor is set or the PC has no PCApplyRichardson): norm, monitor and
 * stopping test of r (unpreconditioned) or B r (preconditioned), x += s B r,
 * r = b - A x; after maxit updates one more norm and test */
static int rich_outer(oksp *k, const double *b, double *x, int *reason, double *rnorm0, double *rnorm) {
    if (k->pc.type == PC_MG && k->pc.mg && !k->monitor) return rich_mg(k, b, x, reason);
    const ocsr *A = k->A;
    const long n = A->n;
    double *r = k->w[0], *z = k->w[1], ttol = 0.0, rn = 0.0;
    ovset(n, 0.0, x);
    ovcopy(n, b, r);
    int i = 0, its = 0;
    *reason = 0;
    for (; i < k->maxit; i++) {
        if (k->normtype == NORM_UNPRECONDITIONED) rn = ovnorm2(n, r);
        else { opc_apply(&k->pc, r, z); rn = ovnorm2(n, z); }
        *rnorm = rn;
        ksp_monitor(k, i, rn);
        *reason = converged(k, i, rn, rnorm0, &ttol);
        if (*reason) break;
        if (k->normtype == NORM_UNPRECONDITIONED) opc_apply(&k->pc, r, z);
        for (long q = 0; q < n; q++) x[q] += k->rich_scale * z[q];
        its++;
        kresid(k, b, x, r);
    }
    if (!*reason) {
        if (k->normtype == NORM_UNPRECONDITIONED) rn = ovnorm2(n, r);
        else { opc_apply(&k->pc, r, z); rn = ovnorm2(n, z); }
        *rnorm = rn;
        ksp_monitor(k, i, rn);
        if (its >= k->maxit) {
            *reason = converged(k, i, rn, rnorm0, &ttol);
            if (!*reason) *reason = R_DIVERGED_ITS;
        }
    }
    return its;
}

int oksp_solve(oksp *k, const double *b, double *x, int *reason, double *rnorm0, double *rnorm) {
    *reason = 0; *rnorm0 = 0.0; *rnorm = 0.0;
    switch (k->type) {
    case KSP_CG: return cg_solve(k, b, x, reason, rnorm0, rnorm);
    case KSP_CHEBYSHEV: *reason = R_CONVERGED_ITS; return cheb_solve(k, b, x);
    case KSP_RICHARDSON:
        if (k->normtype != NORM_NONE) return rich_outer(k, b, x, reason, rnorm0, rnorm);
        *reason = R_CONVERGED_ITS;
        return rich_solve(k, b, x);
    case KSP_PREONLY: opc_apply(&k->pc, b, x); *reason = R_CONVERGED_ITS; return 1;
    case KSP_GMRES: return gmres_solve(k, b, x, reason, rnorm0, rnorm);
    }
    return 0;
}

/* ------------------------------------------------------------------- MG */

/* Galerkin coarse operator A_c = P^T A P */
void oracle_galerkin(const ocsr *A, const ocsr *P, long nc, ocsr *Ac) {
    /* transpose of P */
    long nf = P->n;
    int64_t *tc = (int64_t *)calloc(nc + 1, sizeof(int64_t));
    for (int64_t q = 0; q < P->rp[nf]; q++) tc[P->ci[q] + 1]++;
    for (long c = 0; c < nc; c++) tc[c + 1] += tc[c];
    int32_t *tr = (int32_t *)malloc(sizeof(int32_t) * P->rp[nf]);
    double *tv = (double *)malloc(sizeof(double) * P->rp[nf]);
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * nc);
    memcpy(fill, tc, sizeof(int64_t) * nc);
    for (long i = 0; i < nf; i++)
        for (int64_t q = P->rp[i]; q < P->rp[i + 1]; q++) {
            int64_t s = fill[P->ci[q]]++;
            tr[s] = (int32_t)i; tv[s] = P->v[q];
        }
    double *acc = (double *)calloc(nc, sizeof(double));
    char *mark = (char *)calloc(nc, 1);
    long *list = (long *)malloc(sizeof(long) * nc);
    long cap = nc * 27 + 16;
    Ac->n = nc;
    Ac->rp = (int64_t *)malloc(sizeof(int64_t) * (nc + 1));
    Ac->ci = (int32_t *)malloc(sizeof(int32_t) * cap);
    Ac->v = (double *)malloc(sizeof(double) * cap);
    Ac->rp[0] = 0;
    int64_t q = 0;
    for (long I = 0; I < nc; I++) {
        long nl = 0;
        for (int64_t a = tc[I]; a < tc[I + 1]; a++) {
            long i = tr[a];
            double pi = tv[a];
            for (int64_t b = A->rp[i]; b < A->rp[i + 1]; b++) {
                long j = A->ci[b];
                double aij = A->v[b];
                for (int64_t c = P->rp[j]; c < P->rp[j + 1]; c++) {
                    long J = P->ci[c];
                    if (!mark[J]) { mark[J] = 1; acc[J] = 0.0; list[nl++] = J; }
                    acc[J] += pi * aij * P->v[c];
                }
            }
        }
        /* sort the column list (insertion sort; short lists) */
        for (long u = 1; u < nl; u++) {
            long t = list[u], v = u;
            while (v > 0 && list[v - 1] > t) { list[v] = list[v - 1]; v--; }
            list[v] = t;
        }
        for (long u = 0; u < nl; u++) {
            if (q >= cap) {
                cap *= 2;
                Ac->ci = (int32_t *)realloc(Ac->ci, sizeof(int32_t) * cap);
                Ac->v = (double *)realloc(Ac->v, sizeof(double) * cap);
            }
            Ac->ci[q] = (int32_t)list[u]; Ac->v[q++] = acc[list[u]];
            mark[list[u]] = 0;
        }
        Ac->rp[I + 1] = q;
    }
    free(tc); free(tr); free(tv); free(fill); free(acc); free(mark); free(list);
}

static int parse_ksp(const char *s, int def) {
    if (!s) return def;
    if (!strcmp(s, "cg")) return KSP_CG;
    if (!strcmp(s, "richardson")) return KSP_RICHARDSON;
    if (!strcmp(s, "chebyshev")) return KSP_CHEBYSHEV;
    if (!strcmp(s, "preonly")) return KSP_PREONLY;
    if (!strcmp(s, "gmres")) return KSP_GMRES;
    return -1;
}
static int parse_pc(const char *s, int def) {
    if (!s) return def;
    if (!strcmp(s, "none")) return PC_NONE;
    if (!strcmp(s, "jacobi")) return PC_JACOBI;
    if (!strcmp(s, "sor")) return PC_SOR;
    /* ICC(0) of this symmetric operator is ILU(0) (U = D L^T) up to rounding */
    if (!strcmp(s, "ilu") || !strcmp(s, "icc")) return PC_ILU;
    if (!strcmp(s, "lu") || !strcmp(s, "cholesky") || !strcmp(s, "redundant")) return PC_LU;
    if (!strcmp(s, "mg")) return PC_MG;
    return -1;
}
int oracle_parse_ksp(const char *s, int def) { return parse_ksp(s, def); }
int oracle_parse_pc(const char *s, int def) { return parse_pc(s, def); }

int omg_setup(omg *mg, const ogrid *fine, const oprob *prob, const ocsr *Afine,
              const double *kfine, int nlev, oopts *o, char *err, int errlen) {
    memset(mg, 0, sizeof *mg);
    mg->nlev = nlev;
    mg->g = (ogrid *)calloc(nlev, sizeof(ogrid));
    mg->A = (ocsr *)calloc(nlev, sizeof(ocsr));
    mg->smooth = (oksp *)calloc(nlev, sizeof(oksp));
    mg->x = (double **)calloc(nlev, sizeof(double *));
    mg->b = (double **)calloc(nlev, sizeof(double *));
    mg->r = (double **)calloc(nlev, sizeof(double *));
    mg->P = (ocsr *)calloc(nlev, sizeof(ocsr));
    const char *ct = oopts_get(o, "pc_mg_cycle_type");
    mg->cycles = (ct && !strcmp(ct, "w")) ? 2 : 1;
    mg->galerkin = oopts_bool(o, "pc_mg_galerkin", 0);
    const char *mgtype = oopts_get(o, "pc_mg_type");
    mg->full = mgtype && !strcmp(mgtype, "full");
    mg->kaskade = mgtype && !strcmp(mgtype, "kaskade");
    mg->mcycles = oopts_int(o, "pc_mg_multiplicative_cycles", 1);
    if (mg->mcycles < 1) mg->mcycles = 1;
    if (mgtype && strcmp(mgtype, "multiplicative") && !mg->full && !mg->kaskade) {
        snprintf(err, errlen, "oracle: -pc_mg_type %s not supported", mgtype);
        return 56;
    }
    mg->g[nlev - 1] = *fine;
    mg->nranks = oopts_int(o, "oracle_nranks", 1);
    if (mg->nranks > 16) mg->nranks = 16;
    for (int l = nlev - 1; l > 0; l--)
        if (ogrid_coarsen(&mg->g[l], &mg->g[l - 1])) {
            snprintf(err, errlen, "oracle: grid of %d levels cannot be coarsened", nlev);
            return 63;
        }
    /* ownership of the slowest dimension per simulated rank: fine sizes
     * M/np + (M%np > i); a coarse point I belongs to the owner of fine 2I */
    if (mg->nranks > 1) {
        int L = nlev - 1, np = mg->nranks;
        int ms = fine->dim == 3 ? fine->mz : (fine->dim == 2 ? fine->my : fine->mx);
        mg->split = (int (*)[17])calloc(nlev, sizeof(int[17]));
        mg->split[L][0] = 0;
        for (int q = 0; q < np; q++) mg->split[L][q + 1] = mg->split[L][q] + ms / np + (ms % np > q);
        for (int l = L; l > 0; l--) {
            const ogrid *gc = &mg->g[l - 1];
            int mc = gc->dim == 3 ? gc->mz : (gc->dim == 2 ? gc->my : gc->mx);
            mg->split[l - 1][0] = 0;
            for (int q = 0; q < np; q++) {
                int cnt = 0;
                for (int I = 0; I < mc; I++)
                    if (2 * I >= mg->split[l][q] && 2 * I < mg->split[l][q + 1]) cnt++;
                mg->split[l - 1][q + 1] = mg->split[l - 1][q] + cnt;
            }
        }
    }
    /* operators */
    double *kc_prev = NULL;
    if (kfine) {
        size_t nc = (size_t)(fine->mx - 1) * (fine->my - 1) * (fine->mz - 1);
        kc_prev = (double *)malloc(sizeof(double) * nc);
        memcpy(kc_prev, kfine, sizeof(double) * nc);
    }
    mg->A[nlev - 1] = *Afine;   /* borrowed: not freed by omg_free */
    for (int l = nlev - 1; l > 0; l--) {
        if (mg->galerkin) {
            oq1_build_P(&mg->g[l - 1], &mg->g[l], &mg->P[l]);
            oracle_galerkin(&mg->A[l], &mg->P[l], mg->g[l - 1].n, &mg->A[l - 1]);
        } else if (kfine) {
            const ogrid *gf = &mg->g[l];
            size_t nc = (size_t)((gf->mx - 1) / 2) * ((gf->my - 1) / 2) * ((gf->mz - 1) / 2);
            double *kc = (double *)malloc(sizeof(double) * (nc ? nc : 1));
            oracle_kcell_coarsen(gf, kc_prev, kc);
            oracle_varcoef_jacobian(&mg->g[l - 1], kc, &mg->A[l - 1]);
            free(kc_prev); kc_prev = kc;
        } else {
            oracle_jacobian(&mg->g[l - 1], prob, &mg->A[l - 1]);
        }
    }
    free(kc_prev);
    /* smoothers (PCMGSetLevels defaults: Chebyshev, SOR, 2 its) and coarse */
    int ltype = parse_ksp(oopts_get(o, "mg_levels_ksp_type"), KSP_CHEBYSHEV);
    int lpc = parse_pc(oopts_get(o, "mg_levels_pc_type"), PC_SOR);
    int lits = oopts_int(o, "mg_levels_ksp_max_it", 2);
    double lscale = oopts_real(o, "mg_levels_ksp_richardson_scale", 1.0);
    int ctype = parse_ksp(oopts_get(o, "mg_coarse_ksp_type"), KSP_PREONLY);
    int cpc = parse_pc(oopts_get(o, "mg_coarse_pc_type"), PC_LU);
    int esteig = oopts_int(o, "mg_levels_ksp_chebyshev_esteig_steps", 10);
    if (ltype < 0 || lpc < 0 || lpc == PC_MG || ctype < 0 || cpc < 0 || cpc == PC_MG) {
        snprintf(err, errlen, "oracle: unsupported level smoother options");
        return 56;
    }
    for (int l = 0; l < nlev; l++) {
        long n = mg->g[l].n;
        if (l > 0 || nlev == 1) {   /* PCMGSetLevels: a single level is a smoother level */
            oksp_init(&mg->smooth[l], ltype, &mg->A[l], lpc);
            mg->smooth[l].maxit = lits;
            mg->smooth[l].rich_scale = lscale;
            mg->smooth[l].esteig_steps = esteig < 1 ? 1 : (esteig > 15 ? 15 : esteig);
        } else {
            oksp_init(&mg->smooth[l], ctype, &mg->A[l], cpc);
            mg->smooth[l].maxit = (ctype == KSP_PREONLY) ? 1 : oopts_int(o, "mg_coarse_ksp_max_it", 10000);
        }
        if (mg->nranks > 1 && (lpc == PC_SOR || (l == 0 && cpc == PC_SOR))) {
            /* DMDA split along the slowest dimension only: local SOR blocks */
            const ogrid *gl = &mg->g[l];
            int ms = gl->dim == 3 ? gl->mz : (gl->dim == 2 ? gl->my : gl->mx);
            long plane = gl->n / ms;
            opc *pc = &mg->smooth[l].pc;
            pc->nblk = mg->nranks;
            for (int q = 0; q <= mg->nranks; q++) pc->blk[q] = (long)mg->split[l][q] * plane;
        }
        if ((l > 0 || nlev == 1) && ltype == KSP_CHEBYSHEV) oksp_estimate_eigs(&mg->smooth[l]);
        mg->smooth[l].guess_nonzero = 1;   /* smoothd == smoothu */
        mg->smooth[l].normtype = NORM_NONE;
        if (l < nlev - 1) {
            mg->x[l] = (double *)calloc(n, sizeof(double));
            mg->b[l] = (double *)calloc(n, sizeof(double));
        }
        mg->r[l] = (double *)calloc(n, sizeof(double));
    }
    if (nlev == 1) mg->smooth[0].guess_nonzero = 0;
    return 0;
}

static void mcycle(omg *mg, int l) {
    int reason; double r0, rn;
    if (l == 0) {
        if (mg->smooth[0].type == KSP_PREONLY) opc_apply(&mg->smooth[0].pc, mg->b[0], mg->x[0]);
        else oksp_solve(&mg->smooth[0], mg->b[0], mg->x[0], &reason, &r0, &rn);
        return;
    }
    oksp_solve(&mg->smooth[l], mg->b[l], mg->x[l], &reason, &r0, &rn);
    ocsr_resid(&mg->A[l], mg->b[l], mg->x[l], mg->r[l]);
    if (l == mg->nlev - 1 && mg->rreason && mg->ttol > 0.0) {   /* PCApplyRichardson_MG */
        const double rr = ovnorm2(mg->g[l].n, mg->r[l]);
        if (rr <= mg->ttol) {
            *mg->rreason = rr < mg->atol ? R_CONVERGED_ATOL : R_CONVERGED_RTOL;
            return;
        }
    }
    oq1_restrict(&mg->g[l], &mg->g[l - 1], mg->r[l], mg->b[l - 1]);
    ovset(mg->g[l - 1].n, 0.0, mg->x[l - 1]);
    for (int c = 0; c < mg->cycles; c++) mcycle(mg, l - 1);
    oq1_interp_add(&mg->g[l - 1], &mg->g[l], mg->x[l - 1], mg->x[l]);
    oksp_solve(&mg->smooth[l], mg->b[l], mg->x[l], &reason, &r0, &rn);
}

/* F-cycle: b restricted through every level to the coarsest; the coarsest
 * solved from zero; then for l = 1 .. L: x_l = P x_{l-1} (overwritten, not
 * added) and one cycle on level l from that iterate */
static void fcycle(omg *mg) {
    int L = mg->nlev - 1;
    for (int l = L; l > 0; l--) oq1_restrict(&mg->g[l], &mg->g[l - 1], mg->b[l], mg->b[l - 1]);
    ovset(mg->g[0].n, 0.0, mg->x[0]);
    mcycle(mg, 0);
    for (int l = 1; l <= L; l++) {
        ovset(mg->g[l].n, 0.0, mg->x[l]);
        oq1_interp_add(&mg->g[l - 1], &mg->g[l], mg->x[l - 1], mg->x[l]);
        mcycle(mg, l);
    }
}

/* Kaskade (PETSc's PCMGKCycle_Private): b restricted through every level,
 * the coarsest solved from zero, then upward x_l = P x_{l-1} (overwritten)
 * and one smoother solve on level l from that iterate (no coarse
 * correction) */
static void kcycle(omg *mg) {
    int reason; double r0, rn;
    int L = mg->nlev - 1;
    for (int l = L; l > 0; l--) oq1_restrict(&mg->g[l], &mg->g[l - 1], mg->b[l], mg->b[l - 1]);
    ovset(mg->g[0].n, 0.0, mg->x[0]);
    mcycle(mg, 0);
    for (int l = 1; l <= L; l++) {
        ovset(mg->g[l].n, 0.0, mg->x[l]);
        oq1_interp_add(&mg->g[l - 1], &mg->g[l], mg->x[l - 1], mg->x[l]);
        oksp_solve(&mg->smooth[l], mg->b[l], mg->x[l], &reason, &r0, &rn);
    }
}

/* PCApply_MG: x = 0, then multiplicative: -pc_mg_multiplicative_cycles
 * cycles, each from the last iterate; full: one F-cycle; kaskade: one
 * K-cycle */
void omg_apply(omg *mg, const double *b, double *x) {
    int L = mg->nlev - 1;
    mg->b[L] = (double *)b;
    mg->x[L] = x;
    ovset(mg->g[L].n, 0.0, x);
    if (mg->full && L > 0) fcycle(mg);
    else if (mg->kaskade && L > 0) kcycle(mg);
    else
        for (int c = 0; c < mg->mcycles; c++) mcycle(mg, L);
    mg->b[L] = NULL; mg->x[L] = NULL;
}

void omg_free(omg *mg) {
    for (int l = 0; l < mg->nlev; l++) {
        oksp_free(&mg->smooth[l]);
        if (l < mg->nlev - 1) { free(mg->x[l]); free(mg->b[l]); ocsr_free(&mg->A[l]); }
        free(mg->r[l]);
        if (mg->P[l].v) ocsr_free(&mg->P[l]);
    }
    free(mg->split);
    free(mg->g); free(mg->A); free(mg->smooth); free(mg->x); free(mg->b); free(mg->r); free(mg->P);
    memset(mg, 0, sizeof *mg);
}
