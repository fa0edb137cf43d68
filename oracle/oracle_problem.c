/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * Problems, grids, residual/Jacobian, Q1 transfers, vectors and options.
 * Restates c/ch6/poissonfunctions.c and the problem tables of c/ch6/fish.c.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_internal.h"

const char *oracle_problem_names[PROB_COUNT] = {
    "manupoly", "manuexp", "zero", "manusin", "sinemodes", "varcoef"};

/* ------------------------------------------------------------------ RNG */

/* Philox-4x32-10 (Salmon et al., SC'11), counter (idx lo, idx hi, 0, 0),
 * key (seed lo, seed hi). */
static void philox4x32(uint64_t seed, uint64_t idx, uint32_t out[4]) {
    uint32_t c0 = (uint32_t)idx, c1 = (uint32_t)(idx >> 32), c2 = 0, c3 = 0;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* raw block function with a full 128-bit counter, for known-answer tests */
void oracle_philox4x32_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u01(uint32_t a) { return ((double)a + 0.5) * (1.0 / 4294967296.0); }

/* standard normal by Box-Muller on the first two Philox words */
double oracle_normal(uint64_t seed, uint64_t index) {
    uint32_t o[4];
    philox4x32(seed, index, o);
    double u1 = u01(o[0]), u2 = u01(o[1]);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static double uniform_at(uint64_t seed, uint64_t index) {
    uint32_t o[4];
    philox4x32(seed, index, o);
    return u01(o[2]);
}

/* splitmix64 -> [0,1): start vector of the Chebyshev eigen-estimate */
double oracle_noisy(int64_t index) {
    uint64_t z = (uint64_t)index + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------- problems */

void oprob_init_tables(oprob *p) {
    for (int m = 0; m < NMODES; m++) {
        uint32_t o[4];
        philox4x32(p->modes_seed ^ 0x5EEDull, (uint64_t)m, o);
        p->mp[m] = 1 + (int)(o[0] % 8u);
        p->mq[m] = 1 + (int)(o[1] % 8u);
        p->mr[m] = (p->dim == 3) ? 1 + (int)(o[2] % 8u) : 0;
        p->ma[m] = oracle_normal(p->modes_seed ^ 0xA11Cull, (uint64_t)m);
    }
    p->ksig = 1.0; p->kell = 0.1;
    for (int m = 0; m < NFEAT; m++) {
        for (int d = 0; d < 3; d++)
            p->kw[m][d] = oracle_normal(p->k_seed ^ 0xC0EFull, (uint64_t)(3 * m + d)) / p->kell;
        p->kphi[m] = 2.0 * M_PI * uniform_at(p->k_seed ^ 0xC0EFull, (uint64_t)m);
    }
}

/* g(x) = sigma sqrt(2/M) sum cos(w.x + phi); k = exp(g) */
static double rff_g(const oprob *p, double x, double y, double z, double grad[3]) {
    double s = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, a = p->ksig * sqrt(2.0 / NFEAT);
    for (int m = 0; m < NFEAT; m++) {
        double ph = p->kw[m][0] * x + p->kw[m][1] * y + p->kw[m][2] * z + p->kphi[m];
        s += cos(ph);
        double sn = sin(ph);
        gx -= p->kw[m][0] * sn; gy -= p->kw[m][1] * sn; gz -= p->kw[m][2] * sn;
    }
    if (grad) { grad[0] = a * gx; grad[1] = a * gy; grad[2] = a * gz; }
    return a * s;
}

double oprob_k(const oprob *p, double x, double y, double z) {
    return exp(rff_g(p, x, y, z, NULL));
}

double oracle_varcoef_k(double x, double y, double z, uint64_t seed) {
    oprob p; memset(&p, 0, sizeof p);
    p.dim = 3; p.k_seed = seed;
    oprob_init_tables(&p);
    return oprob_k(&p, x, y, z);
}

/* exact solution = boundary data g (fish.c:15-41) */
double oprob_g(const oprob *p, double x, double y, double z) {
    switch (p->problem) {
    case PROB_MANUPOLY:
        if (p->dim == 1) return x * x * (1.0 - x * x);
        if (p->dim == 2) return x * x * (1.0 - x * x) * y * y * (y * y - 1.0);
        return x * x * (1.0 - x * x) * y * y * (y * y - 1.0) * z * z * (z * z - 1.0);
    case PROB_MANUEXP:
        if (p->dim == 1) return -exp(x);
        if (p->dim == 2) return -x * exp(y);
        return -x * exp(y + z);
    case PROB_ZERO: return 0.0;
    case PROB_MANUSIN:
        if (p->dim == 1) return sin(M_PI * x);
        if (p->dim == 2) return sin(M_PI * x) * sin(2.0 * M_PI * y);
        return sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    case PROB_VARCOEF:
        return sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    case PROB_SINEMODES: {
        /* continuous solution sum a_m/(pi^2 (cx p^2+..)) mode_m */
        double s = 0.0;
        for (int m = 0; m < NMODES; m++) {
            double lam = M_PI * M_PI * (p->cx * p->mp[m] * p->mp[m] + p->cy * p->mq[m] * p->mq[m]
                                        + (p->dim == 3 ? p->cz * p->mr[m] * p->mr[m] : 0.0));
            double md = sin(p->mp[m] * M_PI * x) * sin(p->mq[m] * M_PI * y);
            if (p->dim == 3) md *= sin(p->mr[m] * M_PI * z);
            s += p->ma[m] / lam * md;
        }
        return s;
    }
    }
    return 0.0;
}

/* right-hand side f = -div(c grad u) (fish.c:45-82) */
double oprob_f(const oprob *p, double x, double y, double z) {
    double f = 0.0;
    switch (p->problem) {
    case PROB_MANUPOLY:
        if (p->dim == 1) { f = p->cx * 12.0 * x * x - 2.0; break; }
        if (p->dim == 2) {
            double aa = x * x * (1.0 - x * x), bb = y * y * (y * y - 1.0);
            double ddaa = 2.0 * (1.0 - 6.0 * x * x), ddbb = 2.0 * (6.0 * y * y - 1.0);
            f = -(p->cx * ddaa * bb + p->cy * aa * ddbb);
            break;
        } else {
            double aa = x * x * (1.0 - x * x), bb = y * y * (y * y - 1.0), cc = z * z * (z * z - 1.0);
            double ddaa = 2.0 * (1.0 - 6.0 * x * x), ddbb = 2.0 * (6.0 * y * y - 1.0),
                   ddcc = 2.0 * (6.0 * z * z - 1.0);
            f = -(p->cx * ddaa * bb * cc + p->cy * aa * ddbb * cc + p->cz * aa * bb * ddcc);
            break;
        }
    case PROB_MANUEXP:
        if (p->dim == 1) f = exp(x);
        else if (p->dim == 2) f = x * exp(y);
        else f = 2.0 * x * exp(y + z);
        break;
    case PROB_ZERO: f = 0.0; break;
    case PROB_MANUSIN:
        if (p->dim == 1) f = p->cx * M_PI * M_PI * sin(M_PI * x);
        else if (p->dim == 2)
            f = (p->cx + 4.0 * p->cy) * M_PI * M_PI * sin(M_PI * x) * sin(2.0 * M_PI * y);
        else
            f = (p->cx + p->cy + p->cz) * M_PI * M_PI * sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
        break;
    case PROB_SINEMODES:
        for (int m = 0; m < NMODES; m++) {
            double md = sin(p->mp[m] * M_PI * x) * sin(p->mq[m] * M_PI * y);
            if (p->dim == 3) md *= sin(p->mr[m] * M_PI * z);
            f += p->ma[m] * md;
        }
        break;
    case PROB_VARCOEF: {
        /* -div(k grad u), u = sin sin sin, k = exp(g) */
        double gr[3];
        double k = exp(rff_g(p, x, y, z, gr));
        double sx = sin(M_PI * x), sy = sin(M_PI * y), sz = sin(M_PI * z);
        double cxx = cos(M_PI * x), cyy = cos(M_PI * y), czz = cos(M_PI * z);
        double lap = -3.0 * M_PI * M_PI * sx * sy * sz;
        double du[3] = {M_PI * cxx * sy * sz, M_PI * sx * cyy * sz, M_PI * sx * sy * czz};
        f = -(k * lap + k * (gr[0] * du[0] + gr[1] * du[1] + gr[2] * du[2]));
        break;
    }
    }
    return f;
}

/* f plus the seeded noise at an interior node (node index from coordinates) */
static double f_at_node(const oprob *p, const ogrid *g, int i, int j, int k) {
    double x = i * g->hx, y = j * g->hy, z = k * g->hz;
    double f = oprob_f(p, x, y, z);
    if (p->noise_sigma != 0.0) {
        uint64_t idx = ((uint64_t)k * (uint64_t)g->my + (uint64_t)j) * (uint64_t)g->mx + (uint64_t)i;
        f += p->noise_sigma * oracle_normal(p->noise_seed, idx);
    }
    return f;
}

/* ---------------------------------------------------------------- grids */

void ogrid_set(ogrid *g, int dim, int mx, int my, int mz, double Lx, double Ly, double Lz) {
    g->dim = dim; g->mx = mx; g->my = (dim >= 2) ? my : 1; g->mz = (dim == 3) ? mz : 1;
    g->n = (long)g->mx * g->my * g->mz;
    g->Lx = Lx; g->Ly = Ly; g->Lz = Lz;
    /* DMGetBoundingBox spacing, poissonfunctions.c:74-76 */
    g->hx = Lx / (g->mx - 1);
    g->hy = (dim >= 2) ? Ly / (g->my - 1) : 0.0;
    g->hz = (dim == 3) ? Lz / (g->mz - 1) : 0.0;
}

int ogrid_coarsen(const ogrid *f, ogrid *c) {
    int m[3] = {f->mx, f->my, f->mz};
    for (int d = 0; d < f->dim; d++) {
        if (m[d] < 3 || (m[d] - 1) % 2) return 1;
        m[d] = (m[d] - 1) / 2 + 1;
    }
    ogrid_set(c, f->dim, m[0], m[1], m[2], f->Lx, f->Ly, f->Lz);
    return 0;
}

#define IDX(g, i, j, k) ((((long)(k)) * (g)->my + (j)) * (g)->mx + (i))

/* ------------------------------------------- residual (poissonfunctions.c) */

void oracle_residual(const ogrid *g, const oprob *p, const double *u, double *F) {
    const int mx = g->mx, my = g->my, mz = g->mz;
    if (g->dim == 1) {
        double h = g->hx;
        for (int i = 0; i < mx; i++) {
            double x = i * h;
            if (i == 0 || i == mx - 1) {
                F[i] = u[i] - oprob_g(p, x, 0.0, 0.0);
                F[i] *= p->cx * (2.0 / h);
            } else {
                double ue = (i + 1 == mx - 1) ? oprob_g(p, x + h, 0.0, 0.0) : u[i + 1];
                double uw = (i - 1 == 0) ? oprob_g(p, x - h, 0.0, 0.0) : u[i - 1];
                F[i] = p->cx * (2.0 * u[i] - uw - ue) / h - h * f_at_node(p, g, i, 0, 0);
            }
        }
        return;
    }
    if (g->dim == 2) {
        double hx = g->hx, hy = g->hy, darea = hx * hy;
        double scx = p->cx * hy / hx, scy = p->cy * hx / hy, scdiag = 2.0 * (scx + scy);
        for (int j = 0; j < my; j++) {
            double y = j * hy;
            for (int i = 0; i < mx; i++) {
                double x = i * hx;
                long id = IDX(g, i, j, 0);
                if (i == 0 || i == mx - 1 || j == 0 || j == my - 1) {
                    F[id] = u[id] - oprob_g(p, x, y, 0.0);
                    F[id] *= scdiag;
                } else {
                    double ue = (i + 1 == mx - 1) ? oprob_g(p, x + hx, y, 0.0) : u[id + 1];
                    double uw = (i - 1 == 0) ? oprob_g(p, x - hx, y, 0.0) : u[id - 1];
                    double un = (j + 1 == my - 1) ? oprob_g(p, x, y + hy, 0.0) : u[id + mx];
                    double us = (j - 1 == 0) ? oprob_g(p, x, y - hy, 0.0) : u[id - mx];
                    F[id] = scdiag * u[id] - scx * (uw + ue) - scy * (us + un)
                            - darea * f_at_node(p, g, i, j, 0);
                }
            }
        }
        return;
    }
    double hx = g->hx, hy = g->hy, hz = g->hz, dvol = hx * hy * hz;
    double scx = p->cx * dvol / (hx * hx), scy = p->cy * dvol / (hy * hy), scz = p->cz * dvol / (hz * hz);
    double scdiag = 2.0 * (scx + scy + scz);
    const long sxy = (long)mx * my;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; k++) {
        double z = k * hz;
        for (int j = 0; j < my; j++) {
            double y = j * hy;
            for (int i = 0; i < mx; i++) {
                double x = i * hx;
                long id = IDX(g, i, j, k);
                if (i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) {
                    F[id] = u[id] - oprob_g(p, x, y, z);
                    F[id] *= scdiag;
                } else {
                    double ue = (i + 1 == mx - 1) ? oprob_g(p, x + hx, y, z) : u[id + 1];
                    double uw = (i - 1 == 0) ? oprob_g(p, x - hx, y, z) : u[id - 1];
                    double un = (j + 1 == my - 1) ? oprob_g(p, x, y + hy, z) : u[id + mx];
                    double us = (j - 1 == 0) ? oprob_g(p, x, y - hy, z) : u[id - mx];
                    double uu = (k + 1 == mz - 1) ? oprob_g(p, x, y, z + hz) : u[id + sxy];
                    double ud = (k - 1 == 0) ? oprob_g(p, x, y, z - hz) : u[id - sxy];
                    F[id] = scdiag * u[id] - scx * (uw + ue) - scy * (us + un) - scz * (uu + ud)
                            - dvol * f_at_node(p, g, i, j, k);
                }
            }
        }
    }
}

/* -------------------------------------------- Jacobian (poissonfunctions.c) */

static void csr_alloc(ocsr *A, long n, long nnzmax) {
    A->n = n;
    A->rp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    A->ci = (int32_t *)malloc(sizeof(int32_t) * nnzmax);
    A->v = (double *)malloc(sizeof(double) * nnzmax);
}

void ocsr_free(ocsr *A) {
    free(A->rp); free(A->ci); free(A->v);
    A->rp = NULL; A->ci = NULL; A->v = NULL; A->n = 0;
}

/* rows with diagonal d and couplings -s{x,y,z} to interior neighbours only
 * (poissonfunctions.c:117-258); columns in increasing order              */
void oracle_jacobian(const ogrid *g, const oprob *p, ocsr *A) {
    const int mx = g->mx, my = g->my, mz = g->mz, dim = g->dim;
    double d, sx, sy = 0.0, sz = 0.0;
    if (dim == 1) {
        d = p->cx * 2.0 / g->hx; sx = p->cx / g->hx;
    } else if (dim == 2) {
        sx = p->cx * g->hy / g->hx; sy = p->cy * g->hx / g->hy; d = 2.0 * (sx + sy);
    } else {
        double dvol = g->hx * g->hy * g->hz;
        sx = p->cx * dvol / (g->hx * g->hx); sy = p->cy * dvol / (g->hy * g->hy);
        sz = p->cz * dvol / (g->hz * g->hz); d = 2.0 * (sx + sy + sz);
    }
    csr_alloc(A, g->n, g->n * (2 * dim + 1));
    const long sxy = (long)mx * my;
    /* row pointers first (parallel fill afterwards) */
    A->rp[0] = 0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long id = IDX(g, i, j, k);
                int interior = i > 0 && i < mx - 1 && (dim < 2 || (j > 0 && j < my - 1))
                               && (dim < 3 || (k > 0 && k < mz - 1));
                int cnt = 1;
                if (interior) {
                    cnt += (i - 1 > 0) + (i + 1 < mx - 1);
                    if (dim >= 2) cnt += (j - 1 > 0) + (j + 1 < my - 1);
                    if (dim == 3) cnt += (k - 1 > 0) + (k + 1 < mz - 1);
                }
                A->rp[id + 1] = A->rp[id] + cnt;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long id = IDX(g, i, j, k);
                int64_t q = A->rp[id];
                int interior = i > 0 && i < mx - 1 && (dim < 2 || (j > 0 && j < my - 1))
                               && (dim < 3 || (k > 0 && k < mz - 1));
                if (!interior) { A->ci[q] = (int32_t)id; A->v[q] = d; continue; }
                if (dim == 3 && k - 1 > 0) { A->ci[q] = (int32_t)(id - sxy); A->v[q++] = -sz; }
                if (dim >= 2 && j - 1 > 0) { A->ci[q] = (int32_t)(id - mx); A->v[q++] = -sy; }
                if (i - 1 > 0) { A->ci[q] = (int32_t)(id - 1); A->v[q++] = -sx; }
                A->ci[q] = (int32_t)id; A->v[q++] = d;
                if (i + 1 < mx - 1) { A->ci[q] = (int32_t)(id + 1); A->v[q++] = -sx; }
                if (dim >= 2 && j + 1 < my - 1) { A->ci[q] = (int32_t)(id + mx); A->v[q++] = -sy; }
                if (dim == 3 && k + 1 < mz - 1) { A->ci[q] = (int32_t)(id + sxy); A->v[q++] = -sz; }
            }
}

/* ------------------------------------------------ variable coefficient */

void oracle_kcell_fill(const ogrid *g, const oprob *p, double *kcell) {
    const int cx = g->mx - 1, cy = g->my - 1, cz = g->mz - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < cz; k++)
        for (int j = 0; j < cy; j++)
            for (int i = 0; i < cx; i++)
                kcell[((long)k * cy + j) * cx + i] =
                    oprob_k(p, (i + 0.5) * g->hx, (j + 0.5) * g->hy, (k + 0.5) * g->hz);
}

/* coarse cell = arithmetic mean of the 8 fine cells it contains */
void oracle_kcell_coarsen(const ogrid *f, const double *kf, double *kc) {
    const int fx = f->mx - 1, fy = f->my - 1;
    const int cx = (f->mx - 1) / 2, cy = (f->my - 1) / 2, cz = (f->mz - 1) / 2;
    for (int k = 0; k < cz; k++)
        for (int j = 0; j < cy; j++)
            for (int i = 0; i < cx; i++) {
                double s = 0.0;
                for (int dk = 0; dk < 2; dk++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int di = 0; di < 2; di++)
                            s += kf[((long)(2 * k + dk) * fy + (2 * j + dj)) * fx + (2 * i + di)];
                kc[((long)k * cy + j) * cx + i] = 0.125 * s;
            }
}

/* face coefficients of vertex (i,j,k): mean of the 4 cells sharing each
 * face's edge; missing cells (outside the box) are skipped for the
 * boundary-row diagonal only (interior vertices have all 8 cells).      */
static inline double kc_at(const ogrid *g, const double *kc, int i, int j, int k) {
    return kc[((long)k * (g->my - 1) + j) * (g->mx - 1) + i];
}

static void varcoef_row(const ogrid *g, const double *kc, int i, int j, int k, double *d,
                        double s[6]) {
    /* s = {west, east, south, north, down, up} couplings (positive) */
    double dvol = g->hx * g->hy * g->hz;
    double ax = dvol / (g->hx * g->hx), ay = dvol / (g->hy * g->hy), az = dvol / (g->hz * g->hz);
    int interior = i > 0 && i < g->mx - 1 && j > 0 && j < g->my - 1 && k > 0 && k < g->mz - 1;
    if (!interior) {
        double sum = 0.0; int cnt = 0;
        for (int dk = -1; dk <= 0; dk++)
            for (int dj = -1; dj <= 0; dj++)
                for (int di = -1; di <= 0; di++) {
                    int ci = i + di, cj = j + dj, ck = k + dk;
                    if (ci < 0 || cj < 0 || ck < 0 || ci >= g->mx - 1 || cj >= g->my - 1 || ck >= g->mz - 1)
                        continue;
                    sum += kc_at(g, kc, ci, cj, ck); cnt++;
                }
        *d = 2.0 * (ax + ay + az) * (sum / cnt);
        for (int q = 0; q < 6; q++) s[q] = 0.0;
        return;
    }
    /* x faces: cells (i-1 or i, j-1..j, k-1..k) */
    double kw = 0.25 * (kc_at(g, kc, i - 1, j - 1, k - 1) + kc_at(g, kc, i - 1, j, k - 1)
                        + kc_at(g, kc, i - 1, j - 1, k) + kc_at(g, kc, i - 1, j, k));
    double ke = 0.25 * (kc_at(g, kc, i, j - 1, k - 1) + kc_at(g, kc, i, j, k - 1)
                        + kc_at(g, kc, i, j - 1, k) + kc_at(g, kc, i, j, k));
    double ks = 0.25 * (kc_at(g, kc, i - 1, j - 1, k - 1) + kc_at(g, kc, i, j - 1, k - 1)
                        + kc_at(g, kc, i - 1, j - 1, k) + kc_at(g, kc, i, j - 1, k));
    double kn = 0.25 * (kc_at(g, kc, i - 1, j, k - 1) + kc_at(g, kc, i, j, k - 1)
                        + kc_at(g, kc, i - 1, j, k) + kc_at(g, kc, i, j, k));
    double kd = 0.25 * (kc_at(g, kc, i - 1, j - 1, k - 1) + kc_at(g, kc, i, j - 1, k - 1)
                        + kc_at(g, kc, i - 1, j, k - 1) + kc_at(g, kc, i, j, k - 1));
    double ku = 0.25 * (kc_at(g, kc, i - 1, j - 1, k) + kc_at(g, kc, i, j - 1, k)
                        + kc_at(g, kc, i - 1, j, k) + kc_at(g, kc, i, j, k));
    s[0] = ax * kw; s[1] = ax * ke; s[2] = ay * ks; s[3] = ay * kn; s[4] = az * kd; s[5] = az * ku;
    *d = s[0] + s[1] + s[2] + s[3] + s[4] + s[5];
}

void oracle_varcoef_jacobian(const ogrid *g, const double *kc, ocsr *A) {
    const int mx = g->mx, my = g->my, mz = g->mz;
    const long sxy = (long)mx * my;
    csr_alloc(A, g->n, g->n * 7);
    A->rp[0] = 0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long id = IDX(g, i, j, k);
                int interior = i > 0 && i < mx - 1 && j > 0 && j < my - 1 && k > 0 && k < mz - 1;
                int cnt = 1;
                if (interior)
                    cnt += (i - 1 > 0) + (i + 1 < mx - 1) + (j - 1 > 0) + (j + 1 < my - 1)
                           + (k - 1 > 0) + (k + 1 < mz - 1);
                A->rp[id + 1] = A->rp[id] + cnt;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long id = IDX(g, i, j, k);
                int64_t q = A->rp[id];
                double d, s[6];
                varcoef_row(g, kc, i, j, k, &d, s);
                int interior = i > 0 && i < mx - 1 && j > 0 && j < my - 1 && k > 0 && k < mz - 1;
                if (!interior) { A->ci[q] = (int32_t)id; A->v[q] = d; continue; }
                if (k - 1 > 0) { A->ci[q] = (int32_t)(id - sxy); A->v[q++] = -s[4]; }
                if (j - 1 > 0) { A->ci[q] = (int32_t)(id - mx); A->v[q++] = -s[2]; }
                if (i - 1 > 0) { A->ci[q] = (int32_t)(id - 1); A->v[q++] = -s[0]; }
                A->ci[q] = (int32_t)id; A->v[q++] = d;
                if (i + 1 < mx - 1) { A->ci[q] = (int32_t)(id + 1); A->v[q++] = -s[1]; }
                if (j + 1 < my - 1) { A->ci[q] = (int32_t)(id + mx); A->v[q++] = -s[3]; }
                if (k + 1 < mz - 1) { A->ci[q] = (int32_t)(id + sxy); A->v[q++] = -s[5]; }
            }
}

void oracle_varcoef_residual(const ogrid *g, const oprob *p, const double *kc, const double *u,
                             double *F) {
    const int mx = g->mx, my = g->my, mz = g->mz;
    const long sxy = (long)mx * my;
    double dvol = g->hx * g->hy * g->hz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long id = IDX(g, i, j, k);
                double x = i * g->hx, y = j * g->hy, z = k * g->hz, d, s[6];
                varcoef_row(g, kc, i, j, k, &d, s);
                if (i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) {
                    F[id] = (u[id] - oprob_g(p, x, y, z)) * d;
                    continue;
                }
                double ue = (i + 1 == mx - 1) ? oprob_g(p, x + g->hx, y, z) : u[id + 1];
                double uw = (i - 1 == 0) ? oprob_g(p, x - g->hx, y, z) : u[id - 1];
                double un = (j + 1 == my - 1) ? oprob_g(p, x, y + g->hy, z) : u[id + mx];
                double us = (j - 1 == 0) ? oprob_g(p, x, y - g->hy, z) : u[id - mx];
                double uu = (k + 1 == mz - 1) ? oprob_g(p, x, y, z + g->hz) : u[id + sxy];
                double ud = (k - 1 == 0) ? oprob_g(p, x, y, z - g->hz) : u[id - sxy];
                F[id] = d * u[id] - s[0] * uw - s[1] * ue - s[2] * us - s[3] * un - s[4] * ud
                        - s[5] * uu - dvol * f_at_node(p, g, i, j, k);
            }
}

int oracle_varcoef_apply(int mx, int my, int mz, double Lx, double Ly, double Lz,
                         const double *kcell, const double *x, double *y) {
    ogrid g; ocsr A;
    ogrid_set(&g, 3, mx, my, mz, Lx, Ly, Lz);
    oracle_varcoef_jacobian(&g, kcell, &A);
    ocsr_mult(&A, x, y);
    ocsr_free(&A);
    return 0;
}

/* ---------------------------------------- initial state and exact solution */

/* InitialState (poissonfunctions.c:260-346): zeros or uniform random, then
 * g on the boundary unless gbdry is false */
void oracle_initial_state(const ogrid *g, const oprob *p, int random, int gbdry, double *u) {
    /* random: the splitmix64 stream of the product's VecSetRandom (seed
     * 0x12345678, counter 1, 2, ...); PETSc's own generator is not
     * reproducible here, and the converged u does not depend on u0 */
    for (long q = 0; q < g->n; q++) {
        if (!random) { u[q] = 0.0; continue; }
        uint64_t z = 0x12345678ull + 0x9E3779B97F4A7C15ull * (uint64_t)(q + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        u[q] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    if (!gbdry) return;
    for (int k = 0; k < g->mz; k++)
        for (int j = 0; j < g->my; j++)
            for (int i = 0; i < g->mx; i++) {
                int bd = i == 0 || i == g->mx - 1;
                if (g->dim >= 2) bd = bd || j == 0 || j == g->my - 1;
                if (g->dim == 3) bd = bd || k == 0 || k == g->mz - 1;
                if (bd) u[IDX(g, i, j, k)] = oprob_g(p, i * g->hx, j * g->hy, k * g->hz);
            }
}

/* Form{1,2,3}DUExact (fish.c:288-340) */
void oracle_uexact(const ogrid *g, const oprob *p, double *u) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < g->mz; k++)
        for (int j = 0; j < g->my; j++)
            for (int i = 0; i < g->mx; i++)
                u[IDX(g, i, j, k)] = oprob_g(p, i * g->hx, j * g->hy, k * g->hz);
}

/* ------------------------------------------------------------ Q1 transfers */

/* xf += P xc; P is the tensor-product linear interpolation between
 * vertex-nested grids (fine 2I <-> coarse I), all vertices included */
void oq1_interp_add(const ogrid *c, const ogrid *f, const double *xc, double *xf) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < f->mz; k++)
        for (int j = 0; j < f->my; j++)
            for (int i = 0; i < f->mx; i++) {
                int ic[2], jc[2], kc[2], ni, nj, nk;
                double wi[2], wj[2], wk[2];
                if (i % 2 == 0) { ic[0] = i / 2; wi[0] = 1.0; ni = 1; }
                else { ic[0] = (i - 1) / 2; ic[1] = ic[0] + 1; wi[0] = wi[1] = 0.5; ni = 2; }
                if (f->dim < 2) { jc[0] = 0; wj[0] = 1.0; nj = 1; }
                else if (j % 2 == 0) { jc[0] = j / 2; wj[0] = 1.0; nj = 1; }
                else { jc[0] = (j - 1) / 2; jc[1] = jc[0] + 1; wj[0] = wj[1] = 0.5; nj = 2; }
                if (f->dim < 3) { kc[0] = 0; wk[0] = 1.0; nk = 1; }
                else if (k % 2 == 0) { kc[0] = k / 2; wk[0] = 1.0; nk = 1; }
                else { kc[0] = (k - 1) / 2; kc[1] = kc[0] + 1; wk[0] = wk[1] = 0.5; nk = 2; }
                double s = 0.0;
                for (int c3 = 0; c3 < nk; c3++)
                    for (int c2 = 0; c2 < nj; c2++)
                        for (int c1 = 0; c1 < ni; c1++)
                            s += wk[c3] * wj[c2] * wi[c1] * xc[IDX(c, ic[c1], jc[c2], kc[c3])];
                xf[IDX(f, i, j, k)] += s;
            }
}

/* bc = P^T rf (unscaled full weighting, weights 1 and 1/2 per dimension) */
void oq1_restrict(const ogrid *f, const ogrid *c, const double *rf, double *bc) {
#pragma omp parallel for schedule(static)
    for (int K = 0; K < c->mz; K++)
        for (int J = 0; J < c->my; J++)
            for (int I = 0; I < c->mx; I++) {
                double s = 0.0;
                int dk0 = (f->dim == 3) ? -1 : 0, dk1 = (f->dim == 3) ? 1 : 0;
                int dj0 = (f->dim >= 2) ? -1 : 0, dj1 = (f->dim >= 2) ? 1 : 0;
                for (int dk = dk0; dk <= dk1; dk++) {
                    int k = 2 * K + dk;
                    if (k < 0 || k >= f->mz) continue;
                    double wk = dk ? 0.5 : 1.0;
                    for (int dj = dj0; dj <= dj1; dj++) {
                        int j = 2 * J + dj;
                        if (j < 0 || j >= f->my) continue;
                        double wj = dj ? 0.5 : 1.0;
                        for (int di = -1; di <= 1; di++) {
                            int i = 2 * I + di;
                            if (i < 0 || i >= f->mx) continue;
                            double wi = di ? 0.5 : 1.0;
                            s += wk * wj * wi * rf[IDX(f, i, j, k)];
                        }
                    }
                }
                bc[IDX(c, I, J, K)] = s;
            }
}

/* P as CSR (fine rows, coarse columns), for Galerkin products */
void oq1_build_P(const ogrid *c, const ogrid *f, ocsr *P) {
    csr_alloc(P, f->n, f->n * 8);
    long q = 0;
    P->rp[0] = 0;
    for (int k = 0; k < f->mz; k++)
        for (int j = 0; j < f->my; j++)
            for (int i = 0; i < f->mx; i++) {
                int ic[2], jc[2], kc[2], ni, nj, nk;
                double wi[2], wj[2], wk[2];
                if (i % 2 == 0) { ic[0] = i / 2; wi[0] = 1.0; ni = 1; }
                else { ic[0] = (i - 1) / 2; ic[1] = ic[0] + 1; wi[0] = wi[1] = 0.5; ni = 2; }
                if (f->dim < 2) { jc[0] = 0; wj[0] = 1.0; nj = 1; }
                else if (j % 2 == 0) { jc[0] = j / 2; wj[0] = 1.0; nj = 1; }
                else { jc[0] = (j - 1) / 2; jc[1] = jc[0] + 1; wj[0] = wj[1] = 0.5; nj = 2; }
                if (f->dim < 3) { kc[0] = 0; wk[0] = 1.0; nk = 1; }
                else if (k % 2 == 0) { kc[0] = k / 2; wk[0] = 1.0; nk = 1; }
                else { kc[0] = (k - 1) / 2; kc[1] = kc[0] + 1; wk[0] = wk[1] = 0.5; nk = 2; }
                for (int c3 = 0; c3 < nk; c3++)
                    for (int c2 = 0; c2 < nj; c2++)
                        for (int c1 = 0; c1 < ni; c1++) {
                            P->ci[q] = (int32_t)IDX(c, ic[c1], jc[c2], kc[c3]);
                            P->v[q++] = wk[c3] * wj[c2] * wi[c1];
                        }
                P->rp[IDX(f, i, j, k) + 1] = q;
            }
}

int oracle_interp_add(int dim, int mxc, int myc, int mzc, const double *xc, double *xf) {
    ogrid c, f;
    ogrid_set(&c, dim, mxc, myc, mzc, 1, 1, 1);
    ogrid_set(&f, dim, 2 * (mxc - 1) + 1, 2 * (myc - 1) + 1, 2 * (mzc - 1) + 1, 1, 1, 1);
    oq1_interp_add(&c, &f, xc, xf);
    return 0;
}

int oracle_restrict(int dim, int mxc, int myc, int mzc, const double *rf, double *bc) {
    ogrid c, f;
    ogrid_set(&c, dim, mxc, myc, mzc, 1, 1, 1);
    ogrid_set(&f, dim, 2 * (mxc - 1) + 1, 2 * (myc - 1) + 1, 2 * (mzc - 1) + 1, 1, 1, 1);
    oq1_restrict(&f, &c, rf, bc);
    return 0;
}

int oracle_poisson_apply(int dim, int mx, int my, int mz, double Lx, double Ly, double Lz,
                         double cx, double cy, double cz, const double *x, double *y) {
    ogrid g; oprob p; ocsr A;
    memset(&p, 0, sizeof p);
    p.dim = dim; p.cx = cx; p.cy = cy; p.cz = cz; p.Lx = Lx; p.Ly = Ly; p.Lz = Lz;
    ogrid_set(&g, dim, mx, my, mz, Lx, Ly, Lz);
    oracle_jacobian(&g, &p, &A);
    ocsr_mult(&A, x, y);
    ocsr_free(&A);
    return 0;
}

/* ----------------------------------------------------------------- CSR ops */

void ocsr_mult(const ocsr *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < A->n; i++) {
        double s = 0.0;
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) s += A->v[q] * x[A->ci[q]];
        y[i] = s;
    }
}

void ocsr_resid(const ocsr *A, const double *b, const double *x, double *r) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < A->n; i++) {
        double s = 0.0;
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) s += A->v[q] * x[A->ci[q]];
        r[i] = b[i] - s;
    }
}

static double csr_get(const ocsr *A, long i, long j) {
    for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++)
        if (A->ci[q] == j) return A->v[q];
    return 0.0;
}

int ocsr_is_symmetric(const ocsr *A, double tol) {
    for (long i = 0; i < A->n; i++)
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++)
            if (fabs(A->v[q] - csr_get(A, A->ci[q], i)) > tol) return 0;
    return 1;
}

/* ----------------------------------------------------------------- vectors */

/* deterministic: fixed chunks of 4096 summed in order, whatever the
 * thread count (an OpenMP reduction combines partials in no fixed order) */
#define OV_CHUNK 4096L
double ovdot(long n, const double *x, const double *y) {
    const long nc = (n + OV_CHUNK - 1) / OV_CHUNK;
    if (nc <= 1) {
        double s = 0.0;
        for (long i = 0; i < n; i++) s += x[i] * y[i];
        return s;
    }
    double *part = (double *)malloc(sizeof(double) * nc);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; c++) {
        const long lo = c * OV_CHUNK, hi = lo + OV_CHUNK < n ? lo + OV_CHUNK : n;
        double s = 0.0;
        for (long i = lo; i < hi; i++) s += x[i] * y[i];
        part[c] = s;
    }
    double s = 0.0;
    for (long c = 0; c < nc; c++) s += part[c];
    free(part);
    return s;
}
double ovnorm2(long n, const double *x) { return sqrt(ovdot(n, x, x)); }
double ovnorminf(long n, const double *x) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long i = 0; i < n; i++) m = fabs(x[i]) > m ? fabs(x[i]) : m;
    return m;
}
void ovcopy(long n, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) y[i] = x[i];
}
void ovset(long n, double a, double *x) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) x[i] = a;
}

/* ----------------------------------------------------------------- options */

static int looks_like_value(const char *s) {
    if (s[0] != '-') return 1;
    return (s[1] >= '0' && s[1] <= '9') || s[1] == '.';
}

void oopts_parse(oopts *o, int argc, const char **argv) {
    o->n = 0;
    for (int a = 0; a < argc && o->n < 256; a++) {
        const char *s = argv[a];
        if (s[0] != '-' || looks_like_value(s)) continue;
        o->key[o->n] = s + 1;
        o->val[o->n] = (a + 1 < argc && looks_like_value(argv[a + 1])) ? argv[++a] : "";
        o->used[o->n] = 0;
        o->n++;
    }
}
const char *oopts_get(oopts *o, const char *key) {
    for (int q = o->n - 1; q >= 0; q--)
        if (!strcmp(o->key[q], key)) { o->used[q] = 1; return o->val[q]; }
    return NULL;
}
int oopts_has(oopts *o, const char *key) { return oopts_get(o, key) != NULL; }
int oopts_int(oopts *o, const char *key, int def) {
    const char *v = oopts_get(o, key);
    return (v && *v) ? atoi(v) : def;
}
double oopts_real(oopts *o, const char *key, double def) {
    const char *v = oopts_get(o, key);
    return (v && *v) ? atof(v) : def;
}
int oopts_bool(oopts *o, const char *key, int def) {
    const char *v = oopts_get(o, key);
    if (!v) return def;
    if (!*v) return 1;
    return !(strcmp(v, "false") == 0 || strcmp(v, "0") == 0 || strcmp(v, "no") == 0);
}

void oprintf(char *buf, int len, int *pos, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (buf && *pos < len) {
        int w = vsnprintf(buf + *pos, (size_t)(len - *pos), fmt, ap);
        if (w > 0) *pos += (w < len - *pos) ? w : len - *pos - 1;
    } else if (!buf) {
        vprintf(fmt, ap);
    }
    va_end(ap);
}
