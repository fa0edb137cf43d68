/*
 * problems.h -- exact solutions u = g and right-hand sides f for the fish
 * driver: the reference's manupoly, manuexp, zero (c/ch6/fish.c:15-82) and
 * the benchmark problems added through the same PointwiseFcn signature
 * (c/ch6/fish.c:113):
 *   manusin    u = sin(pi x) [sin(2 pi y) in 2D | sin(pi y) sin(pi z) in 3D]
 *   sinemodes  f = sum_{m<8} a_m prod_d sin(p_md pi x_d), a_m ~ N(0,1),
 *              p_md in 1..8 (seeded); u = its continuous solution
 *   varcoef    -div(k grad u) = f, u = sin sin sin, k = exp(g), g a Gaussian
 *              random field (32 random Fourier features, sigma 1, length 0.1)
 * plus optional seeded N(0, sigma^2) noise added to f at interior nodes.
 * Random numbers: Philox-4x32-10 keyed by the seed, counter = node index.
 */
#ifndef P4G_FISH_PROBLEMS_H
#define P4G_FISH_PROBLEMS_H

#include <stdint.h>

enum { PR_MANUPOLY = 0, PR_MANUEXP, PR_ZERO, PR_MANUSIN, PR_SINEMODES, PR_VARCOEF, PR_COUNT };
extern const char *fish_problem_names[PR_COUNT];

#define FISH_NMODES 8
#define FISH_NFEAT 32

typedef struct {
    int    dim, problem;
    double cx, cy, cz;
    double noise_sigma;
    uint64_t noise_seed, modes_seed, k_seed;
    int    mx, my, mz;          /* finest grid (noise node index) */
    double h[3];
    int    mp[FISH_NMODES], mq[FISH_NMODES], mr[FISH_NMODES];
    double ma[FISH_NMODES];
    double kw[FISH_NFEAT][3], kphi[FISH_NFEAT];
} FishProblem;

void   fish_problem_init(FishProblem *p);
double fish_g(const FishProblem *p, double x, double y, double z);
double fish_f(const FishProblem *p, double x, double y, double z);   /* incl. noise */
double fish_k(const FishProblem *p, double x, double y, double z);
double fish_normal(uint64_t seed, uint64_t index);

#endif
