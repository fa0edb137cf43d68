/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference's structured-grid Poisson path
 * (c/ch6/fish.c with c/ch6/poissonfunctions.c) together with the PETSc
 * solver semantics that path reaches through its options (KSPCG, PCMG,
 * KSPCHEBYSHEV, KSPRICHARDSON, PCJACOBI, PCSOR, PCILU, PCLU), and of the
 * obstacle problem of c/ch12/obstacle.c under SNESVINEWTONRSLS.  It exists to
 * check the HIP product library and to time a CPU baseline; nothing in the
 * product links, loads or calls it.
 *
 * Pinning (how each part is checked, see tests/test_oracle.py):
 *   - discretization (residual, Jacobian):  golden fish.test1/2/6/7 error
 *     lines, |F(u0)|_2 = 0.925546 (fish.test1:1), closed-form eigen-mode
 *     errors for the sine problems, brute-force dense solves on tiny grids;
 *   - CG + PCMG + Chebyshev/SOR:  golden iteration counts fish.test1 (8) and
 *     fish.test3 (4), and the printed errors of fish.test3;
 *   - CG + ILU(0):  fish.test5/6/8 (iterations and trajectory-dependent
 *     errors at the default rtol 1e-5);
 *   - Q1 transfer:  <P x, y> = <x, P^T y> identity and exact interpolation
 *     of multilinear functions;
 *   - V-cycle symmetry <Bx,y> = <x,By>, coarse Cholesky vs dense solve;
 *   - Lanczos eigen-estimate vs closed-form lambda_max(D^-1 A) (upper bound)
 *     and vs the GMRES/Hessenberg estimate (equal for constant diagonal).
 *   - operators formed as matrices on small grids against closed forms
 *     (tests/test_cycles.py, tests/test_oracle.py): the three-level V/W
 *     cycles against the two-grid recursion S (I - P C P^T A) S with the
 *     coarser grid's two-level cycle, the three-level F- and Kaskade cycles
 *     against nested iterations; the two-level Galerkin cycle against
 *     (P^T A P)^-1; the Chebyshev smoother against T_k((theta - D^-1 A) /
 *     delta) / T_k(theta/delta); the SOR and Jacobi Richardson smoothers
 *     against the SSOR and (I - s D^-1 A) powers, rank-local SOR against
 *     block SSOR over the z-slabs; ILU(0)/ICC(0) against a
 *     dense pattern-restricted elimination; the CG norm types against
 *     B = I and B = I/4; DIVERGED_INDEFINITE_PC on an indefinite V-cycle.
 *   - variable-coefficient operator (config "varcoef"): parity unpinned
 *     against the reference, which has constant cx, cy, cz only; pinned by
 *     symmetry, constant-k reduction to the constant-coefficient stencil
 *     and O(h^2) convergence of the manufactured solution.
 */
#ifndef P4G_ORACLE_H
#define P4G_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result of one fish-style run; the same layout as fish_result in the
 * product's fish driver (p4pdes_amd/fish/fish.h) so tests compare fields. */
typedef struct {
    int    dim, mx, my, mz;
    int    its;              /* outer KSP iterations (ksp->its)            */
    int    reason;           /* KSPConvergedReason numbering               */
    int    inner_its;        /* mixed precision: total inner iterations   */
    int    nlevels;
    double fnorm0, fnorm1;   /* |F(u0)|_2, |F(u)|_2                        */
    double rnorm0, rnorm;    /* norm used by the convergence test          */
    double true_relres;      /* |b - A x|_2 / |b|_2 of the linear solve    */
    double errinf, err2h;    /* fish.c:256-276                             */
    double setup_ms, solve_ms;
    double emax_fine;        /* Chebyshev upper bound on the finest level  */
} fish_result;

/* Run the oracle exactly as `fish` would be run with argv (PETSc-style
 * options).  Text that fish would print goes to outbuf (may be NULL).
 * u_out (may be NULL) receives the solution, mx*my*mz doubles, x fastest.
 * Returns 0 on success, a nonzero PETSc-like error code otherwise. */
int oracle_fish_run(int argc, const char **argv, fish_result *res,
                    double *u_out, char *outbuf, int outlen);

/* Result of one obstacle run (c/ch12/obstacle.c); the same layout as
 * obstacle_result in the product's obstacle driver. */
typedef struct {
    int    mx, my;
    int    its;              /* SNES iterations on the final grid          */
    int    reason;           /* SNESConvergedReason numbering              */
    int    last_ksp_its;     /* KSPGetIterationNumber after the last solve */
    int    ksp_its_total;    /* final grid                                 */
    int    stages;           /* 1 + -snes_grid_sequence                    */
    int    stage_its[16];    /* SNES iterations per grid (first 16)      */
    long   active;           /* GetActiveSet count (obstacle.c:201-232)    */
    double fnorm0, fnorm;    /* inactive-set |F|, final grid               */
    double error1, errorinf, areaerr;   /* obstacle.c:180-196             */
    double setup_ms, solve_ms;
} obstacle_result;

/* Run the oracle as `obstacle` would be run with argv; text to outbuf,
 * the final solution (mx*my doubles) to u_out; both may be NULL. */
int oracle_obstacle_run(int argc, const char **argv, obstacle_result *res,
                        double *u_out, char *outbuf, int outlen);

/* Result of one minimal-surface run (c/ch7/minimal.c); the same layout as
 * minimal_result in the product's minimal driver. */
typedef struct {
    int    mx, my;
    int    its;              /* SNES iterations on the final grid          */
    int    reason;           /* SNESConvergedReason                        */
    int    last_ksp_its, ksp_its_total, stages;
    int    stage_its[16];
    double fnorm0, fnorm;    /* |F|, final grid                            */
    double errinf;           /* catenoid, q = -1/2: |u - u_exact|_inf      */
    double area;             /* MSEMonitor's area of the final iterate     */
    double setup_ms, solve_ms;
} minimal_result;

/* Run the oracle as `minimal` would be run with argv; text to outbuf,
 * the final solution (mx*my doubles) to u_out; both may be NULL. */
int oracle_minimal_run(int argc, const char **argv, minimal_result *res,
                       double *u_out, char *outbuf, int outlen);

/* The minimal-surface residual and its colored finite-difference Jacobian
 * (9 coefficients per row) on an mx x my grid of (0,1)^2; problem 0 tent,
 * 1 catenoid. */
void oracle_minimal_residual(int mx, int my, int problem, double q, double H, double c,
                             const double *u, double *F);
void oracle_minimal_fdjac(int mx, int my, int problem, double q, double H, double c,
                          const double *u, double *J9);
/* y = J(u) a by the matrix-free finite difference of -snes_mf_operator */
void oracle_minimal_mffd(int mx, int my, int problem, double q, double H, double c,
                         const double *u, const double *a, double *y);

/* Time `nsolves` solves (after `warmup`) of the linear system fish builds
 * for argv, with `nthreads` OpenMP threads.  Setup is excluded.  Fills
 * times_ms[nsolves] and the result of the last solve. */
int oracle_fish_time(int argc, const char **argv, int nthreads, int warmup,
                     int nsolves, double *times_ms, fish_result *res);

/* As oracle_fish_time, on one set-up (whose wall time goes to *setup_ms),
 * followed, when sample_its > 0, by one solve on a single thread capped at
 * sample_its iterations: its wall time to *sample_ms and its iteration count
 * to *sample_done. */
int oracle_fish_time_sample(int argc, const char **argv, int nthreads, int warmup,
                            int nsolves, double *times_ms, fish_result *res,
                            double *setup_ms, int sample_its, double *sample_ms,
                            int *sample_done);

/* One preconditioner application z = B r (PCApply), one operator
 * application y = A x (MatMult), and F(u0), for the system fish builds
 * from argv.  Arrays have mx*my*mz doubles. */
int oracle_pc_apply(int argc, const char **argv, const double *r, double *z);
int oracle_mat_mult(int argc, const char **argv, const double *x, double *y);
int oracle_rhs(int argc, const char **argv, double *F);

/* Building blocks exported for the unit tests. ----------------------- */

/* Apply the Poisson operator of a dim-D grid (constant cx,cy,cz, rows as
 * Poisson{1,2,3}DJacobianLocal assembles them) to x: y = A x.            */
int oracle_poisson_apply(int dim, int mx, int my, int mz, double Lx, double Ly,
                         double Lz, double cx, double cy, double cz,
                         const double *x, double *y);

/* Q1 transfers between vertex-nested grids (fine m = 2(mc-1)+1 per dim). */
int oracle_interp_add(int dim, int mxc, int myc, int mzc, const double *xc,
                      double *xf);
int oracle_restrict(int dim, int mxc, int myc, int mzc, const double *rf,
                    double *bc);

/* Deterministic start vector used by the Chebyshev eigen-estimate. */
double oracle_noisy(int64_t index);

/* Philox-4x32-10 block function (Salmon et al. 2011), full counter. */
void oracle_philox4x32_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Philox-4x32-10 normal variate for node `index` and `seed` (RHS noise). */
double oracle_normal(uint64_t seed, uint64_t index);

/* Variable-coefficient (cell-centred k) operator on a 3D vertex grid:
 * y = A_k x with the face coefficient the mean of the 4 cells sharing the
 * face's edge; kcell has (mx-1)(my-1)(mz-1) entries, x fastest.          */
int oracle_varcoef_apply(int mx, int my, int mz, double Lx, double Ly,
                         double Lz, const double *kcell, const double *x,
                         double *y);

/* Cell-coefficient random field k = exp(g) used by the "varcoef" problem. */
double oracle_varcoef_k(double x, double y, double z, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
