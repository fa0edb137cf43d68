/*
 * problems.c -- problem table of the fish driver (see problems.h).
 */
#include <math.h>
#include <string.h>
#include "problems.h"

const char *fish_problem_names[PR_COUNT] = {"manupoly", "manuexp", "zero",
                                            "manusin", "sinemodes", "varcoef"};

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of two
 * 32x32->64 multiplies with Weyl-sequence key bumps. */
static void philox(uint64_t key, uint64_t ctr, uint32_t out[4]) {
    uint32_t x[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    for (int round = 0; round < 10; round++) {
        const uint64_t a = (uint64_t)0xD2511F53u * x[0];
        const uint64_t b = (uint64_t)0xCD9E8D57u * x[2];
        const uint32_t y0 = (uint32_t)(b >> 32) ^ x[1] ^ k[0];
        const uint32_t y2 = (uint32_t)(a >> 32) ^ x[3] ^ k[1];
        x[1] = (uint32_t)b;
        x[3] = (uint32_t)a;
        x[0] = y0;
        x[2] = y2;
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    memcpy(out, x, sizeof x);
}

static double unit(uint32_t w) { return ((double)w + 0.5) / 4294967296.0; }

double fish_normal(uint64_t seed, uint64_t index) {
    uint32_t o[4];
    philox(seed, index, o);
    return sqrt(-2.0 * log(unit(o[0]))) * cos(2.0 * M_PI * unit(o[1]));
}

void fish_problem_init(FishProblem *p) {
    for (int m = 0; m < FISH_NMODES; m++) {
        uint32_t o[4];
        philox(p->modes_seed ^ 0x5EEDull, (uint64_t)m, o);
        p->mp[m] = 1 + (int)(o[0] % 8u);
        p->mq[m] = 1 + (int)(o[1] % 8u);
        p->mr[m] = p->dim == 3 ? 1 + (int)(o[2] % 8u) : 0;
        p->ma[m] = fish_normal(p->modes_seed ^ 0xA11Cull, (uint64_t)m);
    }
    const double ell = 0.1;
    for (int m = 0; m < FISH_NFEAT; m++) {
        for (int d = 0; d < 3; d++)
            p->kw[m][d] = fish_normal(p->k_seed ^ 0xC0EFull, (uint64_t)(3 * m + d)) / ell;
        uint32_t o[4];
        philox(p->k_seed ^ 0xC0EFull, (uint64_t)m, o);
        p->kphi[m] = 2.0 * M_PI * unit(o[2]);
    }
}

static double field(const FishProblem *p, double x, double y, double z, double grad[3]) {
    const double amp = sqrt(2.0 / FISH_NFEAT);   /* sigma = 1 */
    double s = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int m = 0; m < FISH_NFEAT; m++) {
        const double ph = p->kw[m][0] * x + p->kw[m][1] * y + p->kw[m][2] * z + p->kphi[m];
        s += cos(ph);
        const double sn = sin(ph);
        gx -= p->kw[m][0] * sn;
        gy -= p->kw[m][1] * sn;
        gz -= p->kw[m][2] * sn;
    }
    if (grad) { grad[0] = amp * gx; grad[1] = amp * gy; grad[2] = amp * gz; }
    return amp * s;
}

double fish_k(const FishProblem *p, double x, double y, double z) {
    return exp(field(p, x, y, z, NULL));
}

static double mode_lambda(const FishProblem *p, int m) {
    double l = p->cx * p->mp[m] * p->mp[m] + p->cy * p->mq[m] * p->mq[m];
    if (p->dim == 3) l += p->cz * p->mr[m] * p->mr[m];
    return M_PI * M_PI * l;
}

static double mode(const FishProblem *p, int m, double x, double y, double z) {
    double v = sin(p->mp[m] * M_PI * x) * sin(p->mq[m] * M_PI * y);
    return p->dim == 3 ? v * sin(p->mr[m] * M_PI * z) : v;
}

double fish_g(const FishProblem *p, double x, double y, double z) {
    switch (p->problem) {
    case PR_MANUPOLY: {
        const double a = x * x * (1.0 - x * x);
        if (p->dim == 1) return a;
        const double b = y * y * (y * y - 1.0);
        if (p->dim == 2) return a * b;
        return a * b * z * z * (z * z - 1.0);
    }
    case PR_MANUEXP:
        return p->dim == 1 ? -exp(x) : (p->dim == 2 ? -x * exp(y) : -x * exp(y + z));
    case PR_MANUSIN:
        if (p->dim == 1) return sin(M_PI * x);
        if (p->dim == 2) return sin(M_PI * x) * sin(2.0 * M_PI * y);
        return sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    case PR_VARCOEF:
        return sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    case PR_SINEMODES: {
        double s = 0.0;
        for (int m = 0; m < FISH_NMODES; m++) s += p->ma[m] / mode_lambda(p, m) * mode(p, m, x, y, z);
        return s;
    }
    default:
        return 0.0;
    }
}

static double f_clean(const FishProblem *p, double x, double y, double z) {
    switch (p->problem) {
    case PR_MANUPOLY: {
        if (p->dim == 1) return p->cx * 12.0 * x * x - 2.0;
        const double a = x * x * (1.0 - x * x), b = y * y * (y * y - 1.0);
        const double a2 = 2.0 * (1.0 - 6.0 * x * x), b2 = 2.0 * (6.0 * y * y - 1.0);
        if (p->dim == 2) return -(p->cx * a2 * b + p->cy * a * b2);
        const double c = z * z * (z * z - 1.0), c2 = 2.0 * (6.0 * z * z - 1.0);
        return -(p->cx * a2 * b * c + p->cy * a * b2 * c + p->cz * a * b * c2);
    }
    case PR_MANUEXP:
        return p->dim == 1 ? exp(x) : (p->dim == 2 ? x * exp(y) : 2.0 * x * exp(y + z));
    case PR_MANUSIN:
        if (p->dim == 1) return p->cx * M_PI * M_PI * sin(M_PI * x);
        if (p->dim == 2) return (p->cx + 4.0 * p->cy) * M_PI * M_PI * sin(M_PI * x) * sin(2.0 * M_PI * y);
        return (p->cx + p->cy + p->cz) * M_PI * M_PI * sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    case PR_SINEMODES: {
        double s = 0.0;
        for (int m = 0; m < FISH_NMODES; m++) s += p->ma[m] * mode(p, m, x, y, z);
        return s;
    }
    case PR_VARCOEF: {
        double gr[3];
        const double k = exp(field(p, x, y, z, gr));
        const double sx = sin(M_PI * x), sy = sin(M_PI * y), sz = sin(M_PI * z);
        const double lap = -3.0 * M_PI * M_PI * sx * sy * sz;
        const double du[3] = {M_PI * cos(M_PI * x) * sy * sz, M_PI * sx * cos(M_PI * y) * sz,
                              M_PI * sx * sy * cos(M_PI * z)};
        return -(k * lap + k * (gr[0] * du[0] + gr[1] * du[1] + gr[2] * du[2]));
    }
    default:
        return 0.0;
    }
}

double fish_f(const FishProblem *p, double x, double y, double z) {
    double f = f_clean(p, x, y, z);
    if (p->noise_sigma != 0.0) {
        const long i = lround(x / p->h[0]);
        const long j = p->dim > 1 ? lround(y / p->h[1]) : 0;
        const long k = p->dim > 2 ? lround(z / p->h[2]) : 0;
        const uint64_t node = ((uint64_t)k * (uint64_t)p->my + (uint64_t)j) * (uint64_t)p->mx + (uint64_t)i;
        f += p->noise_sigma * fish_normal(p->noise_seed, node);
    }
    return f;
}
