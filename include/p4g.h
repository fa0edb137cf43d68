/*
 * p4g.h -- C ABI of the MI355X (gfx950) CG + geometric-multigrid Poisson
 * solver.  It stands in for the PETSc objects the reference path creates in
 * c/ch6/fish.c:197-239 when run as `-ksp_type cg -pc_type mg`:
 *
 *   DMDACreate{1,2,3}d + DMSetFromOptions + DMDASetUniformCoordinates
 *        -> p4g_da_create, p4g_da_set_from_options
 *   DMDASNESSetFunctionLocal / DMDASNESSetJacobianLocal
 *        -> p4g_da_set_function_local / p4g_da_set_jacobian_local
 *   SNESCreate/SNESSetDM/SNESSetType(KSPONLY)/KSPSetType(CG)/SNESSetFromOptions
 *        -> p4g_solver_create (PETSc-style option string)
 *   SNESSolve(snes, NULL, u)   -> p4g_snes_solve
 *   KSPSolve / PCApply / MatMult on the assembled operator
 *        -> p4g_ksp_solve[_device], p4g_pc_apply[_device], p4g_mat_mult[_device]
 *
 * Conventions: every entry point returns 0 or a PETSc error number
 * (PETSC_ERR_* in p4g_petsc.h) and leaves a message for p4g_last_error()
 * (thread-local).  No C++ exception crosses the ABI.  The caller owns host
 * arrays; the library owns device memory.  Host arrays are mx*my*mz doubles,
 * x fastest (the PETSc natural ordering of a one-process DMDA).  One solver
 * handle per host thread; each solver launches on its own HIP stream.
 * Device pointers passed to *_device entry points must be allocated on the
 * current HIP device (p4g_device_malloc or any hipMalloc).
 */
#ifndef P4G_H
#define P4G_H

#include <stddef.h>
#include <stdint.h>
/* A program that also includes PETSc's own headers (a PCSHELL/MATSHELL
 * binding inside a real PETSc build) defines P4G_HOSTED_BY_PETSC: the
 * PETSc-shaped host surface is then left out and grids are p4g_DA handles;
 * the callback registration needs that surface and is left out too. */
#ifndef P4G_HOSTED_BY_PETSC
#include "p4g_petsc.h"
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct p4g_solver_s *p4g_Solver;
typedef struct p4g_da_s *p4g_DA;   /* the DM of p4g_petsc.h */

/* KSPConvergedReason numbering (PETSc) */
#define P4G_CONVERGED_RTOL 2
#define P4G_CONVERGED_ATOL 3
#define P4G_CONVERGED_ITS 4
#define P4G_DIVERGED_NULL (-2)
#define P4G_DIVERGED_ITS (-3)
#define P4G_DIVERGED_DTOL (-4)
#define P4G_DIVERGED_BREAKDOWN (-5)
#define P4G_DIVERGED_INDEFINITE_PC (-8)
#define P4G_DIVERGED_NANORINF (-9)
#define P4G_DIVERGED_INDEFINITE_MAT (-10)

typedef struct {
    int    its;            /* outer iterations (PETSc ksp->its)                   */
    int    reason;         /* KSPConvergedReason                                  */
    int    inner_its;      /* mixed precision: FP64 residual replacements made    */
    int    nlevels;        /* multigrid levels                                    */
    double rnorm0, rnorm;  /* norm used by the convergence test, first and last  */
    double true_relres;    /* |b - A x|_2 / |b|_2 computed in FP64 after the solve */
    double fnorm0, fnorm1; /* p4g_snes_solve: |F(u0)|_2 and |F(u)|_2             */
    double setup_ms;       /* last set-up (callbacks, hierarchy, eigen-estimates) */
    double solve_ms;       /* last KSP solve, device time (HIP events)           */
    double emax_fine;      /* Chebyshev upper bound on the finest level          */
    double history[64];    /* convergence-test norms of iterations 0..63         */
} p4g_Report;

/* ---------------------------------------------------------------- grids */

/* DMDACreate{1,2,3}d(DM_BOUNDARY_NONE, STAR, dof 1, width 1) with a global
 * size of mx (x my (x mz)) vertices on [0,Lx]x[0,Ly]x[0,Lz]
 * (fish.c:197-220; the reference starts from 3 points per dimension). */
int p4g_da_create(int dim, int mx, int my, int mz, double Lx, double Ly, double Lz, p4g_DA *da);
/* DMSetFromOptions: -da_grid_x/y/z N, -da_refine R (m <- 2(m-1)+1, R
 * times); the refinement count is the default number of multigrid levels
 * minus one, as with PETSc's PCMG on a refined DMDA. */
int p4g_da_set_from_options(p4g_DA da, const char *options);
int p4g_da_get_sizes(p4g_DA da, int *dim, int *mx, int *my, int *mz, int *refine);
#ifndef P4G_HOSTED_BY_PETSC
/* DMDASNESSetFunctionLocal(da, INSERT_VALUES, fn, ctx)  (fish.c:225-226) */
int p4g_da_set_function_local(p4g_DA da, DMDASNESFunctionFn *fn, void *ctx);
/* DMDASNESSetJacobianLocal(da, fn, ctx)                 (fish.c:227-228) */
int p4g_da_set_jacobian_local(p4g_DA da, DMDASNESJacobianFn *fn, void *ctx);
#endif
/* Variable-coefficient operator -div(k grad u) with k given per cell,
 * (mx-1)(my-1)(mz-1) values (3D only), face coefficient = mean of the 4
 * cells sharing the face's edge, coarse k = mean of the 8 fine cells.
 * Replaces the Jacobian callback's operator (the reference's callbacks
 * have constant cx,cy,cz only, poissonfunctions.h:44-47). */
int p4g_da_set_cell_coefficient(p4g_DA da, const double *kcell);
int p4g_da_destroy(p4g_DA *da);

/* --------------------------------------------------------------- solver */

/* SNES KSPONLY + KSPCG + options (fish.c:223-234).  Supported options:
 *   -ksp_type cg|gmres|richardson (gmres: FP64, left preconditioning,
 *       -ksp_gmres_restart n; richardson: -ksp_richardson_scale s),
 *   -ksp_rtol, -ksp_atol, -ksp_divtol, -ksp_max_it, -ksp_monitor,
 *   -ksp_norm_type preconditioned|unpreconditioned|natural,
 *   -pc_type mg|jacobi|none|ilu|icc (ilu/icc: ILU(0) = IC(0) of this
 *       operator, PETSc's serial default PC; FP64, constant coefficients),
 *   -pc_mg_levels L, -pc_mg_cycle_type v|w,
 *   -pc_mg_type multiplicative|full|kaskade, -pc_mg_multiplicative_cycles n,
 *   -pc_mg_galerkin [both|pmat|mat|none] (P^T A P coarse operators, FP64),
 *   -mg_levels_ksp_type chebyshev|richardson,
 *   -mg_levels_pc_type sor|jacobi|none|ilu|icc (sor, PETSc's default: the
 *       symmetric omega = 1 sweep, one iteration; -p4g_sor_blocks N sweeps
 *       N ranks' z-blocks locally as PETSc's parallel SOR does; FP64,
 *       constant coefficients),
 *   -mg_levels_ksp_max_it n, -mg_levels_ksp_richardson_scale s,
 *   -mg_levels_ksp_chebyshev_esteig_steps n,
 *   -mg_coarse_ksp_type preonly with -mg_coarse_pc_type lu|cholesky|redundant,
 *       or -mg_coarse_ksp_type cg -mg_coarse_ksp_convergence_test skip
 *       -mg_coarse_ksp_max_it n with -mg_coarse_pc_type icc|ilu|jacobi|none
 *       (a fixed-count PCG coarse solve),
 *   -ksp_view_mat|-ksp_view_rhs|-ksp_view_solution binary:<file>,
 *   -p4g_precision double|mixed  (mixed: PCG with FP32 direction, residual
 *       and preconditioner vectors and an FP32 V-cycle, an FP64 iterate
 *       (x += alpha p in FP64), FP64 dot accumulation and coarsest solve, and
 *       FP64 residual replacement: b - A x is recomputed in FP64 whenever the
 *       iterated residual has fallen by -p4g_reliable_delta (default 0.1);
 *       converged when the FP64 |b-Ax|_2 <= max(rtol |b|_2, atol)),
 *   -p4g_reliable_delta d,
 *   -p4g_fp32_diff_form 1|0 (default 1): FP32 interior rows evaluated as
 *       sums of differences s (u_c - u_nb), the same row when
 *       d = 2(sx+sy+sz), with far less rounding on smooth vectors,
 *   -p4g_mixed_x64 1|0 (default 1): 0 keeps the correction since the last
 *       replacement in FP32 and folds it into the FP64 iterate at each
 *       replacement,
 *   -p4g_x_defer d (0..15; default 6 in mixed precision; in the FP64 CG
 *       of p4g_ksp_solve / p4g_snes_solve as many as 16 GiB of directions
 *       hold, at most 15, 0 if fewer than 3; 0 in the Newton solves): the
 *       FP64 iterate updates x += alpha p of up to d iterations are
 *       deferred, their directions kept (d more vectors), and applied in
 *       iteration order (the same FP64 rounding as applying each at once):
 *       in mixed precision at the next residual replacement (one is forced
 *       when d are pending), in FP64 when d are pending and at the end of
 *       the solve; 0 applies each in the next direction pass,
 *   -p4g_fuse_cg auto|1|0 (default auto): the CG residual update fused into
 *       the first kernel of the multigrid cycle and the direction update
 *       with the operator apply; auto fuses the residual update when one
 *       fine-level vector is larger than 256 MiB and the direction update
 *       from 768 MiB (below that the separate streaming passes are
 *       faster); -p4g_fuse_r and -p4g_fuse_dir (auto|1|0, default: the
 *       value of -p4g_fuse_cg) set the two fusions separately,
 *   -p4g_lds auto|1|0 (default auto: 3D levels of 400+ points per side):
 *       stencil passes with their neighbour rows staged in LDS,
 *   -p4g_resid_restrict auto|1|0 (default auto: where -p4g_lds applies):
 *       the cycle's residual and restriction in one pass,
 *       -p4g_rr_rows 8|4 its coarse rows per block, -p4g_rr_planes n its
 *       coarse planes per block (default 0: by grid size),
 *   -p4g_post_fuse auto|1|0 (auto and 1: levels where -p4g_lds applies,
 *       with two or more Chebyshev steps): the prolongation fused with the
 *       first post-smoothing step,
 *   -p4g_sor_pipeline 1|0 (default 1): SOR levels swept by pipelined
 *       strips, one launch per direction (else one per hyperplane).
 * Options and combinations the device path does not implement are rejected
 * with PETSC_ERR_SUP and a message naming them, never silently replaced:
 * e.g. -pc_type lu|sor, -pc_type asm outside p4g_vi_solve, -pc_mg_type
 * additive, ILU with fill, SOR with omega != 1 or one-directional sweeps,
 * -mg_levels_ksp_norm_type, and SOR, ILU, GMRES, Richardson, Galerkin or
 * kaskade together with -p4g_precision mixed. */
int p4g_solver_create(p4g_DA da, const char *options, p4g_Solver *s);
/* PCSetUp_MG equivalent: Jacobian callbacks per level, stencil extraction,
 * device hierarchy, Chebyshev eigen-estimates, coarse Cholesky. */
int p4g_solver_setup(p4g_Solver s);
/* SNESSolve(snes, NULL, u), KSPONLY (fish.c:239): F(u0) by the residual
 * callback, J y = F(u0) by PCG from y = 0, u <- u0 - y.  u is host. */
int p4g_snes_solve(p4g_Solver s, double *u, p4g_Report *rep);
/* KSPSolve(ksp, b, x) from x = 0, host arrays (copies over PCIe). */
int p4g_ksp_solve(p4g_Solver s, const double *b, double *x, p4g_Report *rep);
/* KSPSolve with device-resident b and x (FP64, n doubles each). */
int p4g_ksp_solve_device(p4g_Solver s, const double *d_b, double *d_x, p4g_Report *rep);
/* PCApply: z = B r, one multigrid cycle (PCShellSetApply-compatible). */
int p4g_pc_apply(p4g_Solver s, const double *r, double *z);
int p4g_pc_apply_device(p4g_Solver s, const double *d_r, double *d_z);
/* p4g_pc_apply_device without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory at the call
 * (code under HIP graph capture, tight serving loops).  The same cycle, its
 * work arrays (every level's vectors and the block partial sums) carved from
 * one 128 MiB device array the library holds on each GPU until the process
 * ends: calls on one GPU share it, so calls that overlap on one GPU race.
 * The solver must be FP64 and already set up (p4g_solver_setup:
 * PETSC_ERR_ORDER otherwise); a hierarchy whose work arrays exceed 128 MiB
 * (FP64 V(1,1) hierarchies above about 129^3 or 2049^2) is refused with
 * PETSC_ERR_ARG_OUTOFRANGE. */
int p4g_pc_apply_device_noalloc(p4g_Solver s, const double *d_r, double *d_z);
/* MatMult with the finest-level operator: y = A x. */
int p4g_mat_mult(p4g_Solver s, const double *x, double *y);
int p4g_mat_mult_device(p4g_Solver s, const double *d_x, double *d_y);
int p4g_solver_get_report(p4g_Solver s, p4g_Report *rep);
/* the residual norms of the last solve's convergence test, iteration 0 on
 * (what -ksp_monitor prints; KSPGetResidualHistory): the first min(cap, *n)
 * into h, their number into *n */
int p4g_solver_get_history(p4g_Solver s, double *h, int cap, int *n);
int p4g_solver_destroy(p4g_Solver *s);

/* --------------------------------------------------------- measurement */

/* Kernel classes timed with HIP events on the solver's stream while
 * timing is on (off by default): one event pair around every launch of
 * the selected class. */
enum {
    P4G_K_STENCIL_DOT = 0, /* w = A p fused with (p,w), the CG operator apply  */
    P4G_K_CG_UPDATE,       /* p = z + b p, x += a p (or x and r together)      */
    P4G_K_RESID_ZG,        /* r = b - A (c b): zero-guess smoothing + residual */
    P4G_K_RESTRICT,        /* b_c = P^T r                                      */
    P4G_K_PROLONG,         /* x = c b + P x_c (with the first post-smoothing
                              step fused where the solver fuses it)            */
    P4G_K_SMOOTH,          /* x <- x + s D^-1 (b - A x) (Chebyshev step)       */
    P4G_K_COARSE,          /* coarsest-grid Cholesky solve                     */
    P4G_K_REDUCE,          /* finalisation of fused dot products               */
    P4G_K_APPLY,           /* y = A x, plain operator apply (MatMult)          */
    P4G_K_CG_RESID,        /* r -= a w, the CG residual update on its own      */
    P4G_K_SMOOTH_ZG,       /* a smoothing step from a zero guess fused with the
                              CG residual update (r = r - a w formed and stored) */
    P4G_K_SMOOTH_LAST,     /* the cycle's last smoothing step, with its dots    */
    P4G_K_OTHER,
    P4G_K_COUNT
};
/* on: 0 off, 1 every launch of the class, 2 only its launches on the finest
 * level (one event pair per cycle instead of one per level) */
int p4g_solver_set_timing(p4g_Solver s, int kernel_class, int on);
/* launches, summed device ms, and algorithmic bytes moved, for the class
 * (optionally only launches on the finest level: fine_only = 1) */
int p4g_solver_kernel_stats(p4g_Solver s, int kernel_class, int fine_only, long *launches,
                            double *ms, double *bytes);
int p4g_solver_reset_stats(p4g_Solver s);

/* -------------------------------------------------------------- serving */

/* A continuous-batching solve server (no counterpart in the reference,
 * whose fish.c solves one problem per process).  Requests are KSP solves
 * A x = b with the operator hierarchy of a set-up solver (a V(1,1)
 * Chebyshev- or Richardson-Jacobi multigrid on a 2D or 3D constant-
 * coefficient grid); each is run with the mixed-precision PCG of
 * -p4g_precision mixed to the solver's -ksp_rtol.  A worker thread admits
 * requests in arrival order when their device footprint fits the HBM
 * budget, advances all active requests one iteration per round with
 * batched launches (equal-sized multigrid levels of different requests
 * share a launch; the V-cycle launch sequence is replayed from a hipGraph
 * until the active set changes) and releases a request's memory when it
 * finishes.  b and x are host arrays owned by the caller; both must stay
 * valid until the request has finished.
 * Options: -p4g_serve_graph 1|0, -p4g_serve_graph_after k (default 32:
 * a round is captured in a hipGraph once the active set has held k rounds),
 * -p4g_serve_max_slots n (default 64), -p4g_serve_staging_mb m (default 1024:
 * pinned host staging for right-hand sides and results, allocated with the
 * server).  b is copied at submit; x is written by a completion thread before
 * the request is reported done. */
/* ------------------------------------------- variational inequalities */

/* SNESSolve with SNESVINEWTONRSLS and SNESVISetComputeVariableBounds
 * (c/ch12/obstacle.c:118-121,130-134): the reduced-space Newton method for
 * xl <= x <= xu, F(x) of the grid's residual callback, the Jacobian of its
 * Jacobian callback.  Per Newton step the active set {x <= xl + 1e-8,
 * F > 0} u {x >= xu - 1e-8, F < 0}, the inactive-set system J_II y = F_I
 * by the solver's KSP (CG; -pc_type mg uses Galerkin coarse operators on
 * coarse inactive sets, as PETSc does under RSLS), the backtracking (bt,
 * alpha 0) or basic line search with trial points projected onto the
 * bounds, and SNESConvergedDefault on |F| over the inactive set.  Options
 * (from p4g_solver_create): -snes_rtol 1e-8, -snes_atol 1e-50, -snes_stol
 * 1e-8, -snes_max_it 50, -snes_divergence_tolerance 1e4,
 * -snes_linesearch_type bt|basic, -snes_type vinewtonrsls.  xu may be NULL
 * (+infinity).  x: initial iterate in, solution out; F (may be NULL): the
 * residual at the solution (SNESGetFunction).  Host arrays of the grid's
 * size.  The solver's fine operator stays the last reduced one afterwards:
 * use a solver for VI solves only.
 * -snes_type vinewtonssls (SNESVINEWTONSSLS, obstacle.test4 of
 * c/ch12/makefile:26-27): the semismooth Newton method on the Fischer-
 * Burmeister function phi = Phi(x - xl, F) (xu must be NULL), per Newton
 * step the semismooth Jacobian diag(Db) J + diag(Da), the solver's KSP with
 * PC ilu (default) or none, the bt line search on |phi| with projected
 * trial points; fnorm0/fnorm are then |phi| and active counts {x <= xl +
 * 1e-8, F > 0} at the solution. */
typedef struct {
    int    its;            /* Newton iterations (snes->iter)                   */
    int    reason;         /* SNESConvergedReason (3 CONVERGED_FNORM_RELATIVE) */
    int    last_ksp_its;   /* iterations of the last linear solve             */
    int    ksp_its_total;
    long   active;         /* size of the last active set                      */
    double fnorm0, fnorm;  /* |F| on the inactive set, first and last          */
    double setup_ms;       /* the solver's set-up (levels from the callbacks), first call */
    double solve_ms;       /* host wall time of the solve without the set-up   */
} p4g_VIReport;
int p4g_vi_solve(p4g_Solver s, const double *xl, const double *xu, double *x, double *F,
                 p4g_VIReport *rep);
/* *nf norms fnorm[i], i = 0..its (the SNES monitor values), and *nk
 * linear solves, Newton step i's ksp_its[i] and ksp_reason[i] (a step that
 * ended the solve in failure has a linear solve but no norm).  Copies the
 * first min(cap, count) entries of each; arrays may be NULL. */
int p4g_vi_get_history(p4g_Solver s, double *fnorm, int *ksp_its, int *ksp_reason, int cap, int *nf,
                       int *nk);

/* ------------------------------------------------- minimal surface */

/* SNESSolve with SNESNEWTONLS and -snes_fd_color (or -snes_mf_operator) for c/ch7/minimal.c's
 * FormFunctionLocal (minimal.c:230-301): the minimal-surface equation
 * -div((1 + |grad u|^2)^q grad u) = 0 on a 2D grid of the unit square
 * (DMDASetUniformCoordinates(0,1,0,1)), u = g on the boundary.  The device
 * evaluates that residual itself; gb holds g at the grid's nodes (only the
 * boundary entries are read), and when the grid has a residual callback
 * (p4g_da_set_function_local) it must agree with the device residual at
 * the initial iterate (else PETSC_ERR_ARG_WRONG).  Per Newton step the
 * Jacobian by colored finite differences (PETSc's "wp" step), the solver's
 * KSP (GMRES unless -ksp_type cg; -ksp_gmres_restart) with PC ILU(0)
 * (default), mg (each coarser level: the finite-difference Jacobian at the
 * injected iterate; coarse band LU) or none, the bt (alpha 1e-4) or basic
 * line search, SNESConvergedDefault.  -snes_mf_operator: the Krylov
 * operator is MatMFFD at the iterate ("wp" step h = 1.49e-8 sqrt(1 + |x|) /
 * |a|), the grid's Jacobian callback (minimal.c registers the Poisson
 * Jacobian) the preconditioning matrix on every level; with neither option
 * that callback's matrix is the Newton Jacobian; -snes_mf is refused.  SNES
 * options as for p4g_vi_solve but -snes_type newtonls; -mat_is_symmetric
 * tol checks each Jacobian.  x:
 * initial iterate in, solution out (host, the grid's size).  mon (may be
 * NULL) is called with the host iterate after the initial residual and
 * after each Newton step, as SNESMonitorSet's monitors are; a nonzero
 * return ends the solve with that error.  Replaces the DMDASNESSetFunction-
 * Local + SNESSolve pair of minimal.c:150-177. */
typedef struct {
    int    its, reason, last_ksp_its, ksp_its_total;
    double fnorm0, fnorm;
    double setup_ms, solve_ms;
} p4g_MSReport;
typedef int (*p4g_MSMonitorFn)(void *ctx, int it, double fnorm, const double *x);
int p4g_ms_solve(p4g_Solver s, double q, const double *gb, double *x, p4g_MSMonitorFn mon, void *monctx,
                 p4g_MSReport *rep);
/* As p4g_vi_get_history, plus jac_sym[i]: MatIsSymmetric of Newton step
 * i's Jacobian under -mat_is_symmetric (1 symmetric, 0 not, -1 unchecked). */
int p4g_ms_get_history(p4g_Solver s, double *fnorm, int *ksp_its, int *ksp_reason, int *jac_sym, int cap,
                       int *nf, int *nk);

typedef struct p4g_server_s *p4g_Server;

typedef struct {
    long   submitted, completed, rejected, active, queued, rounds, graph_captures;
    double p50_ms, p99_ms, max_ms, mean_ms; /* latency: submit -> result on host */
    double queue_p50_ms;                    /* submit -> admission                */
    double wall_s;                          /* first submit -> last completion    */
    double dof_per_s;                       /* DOF of completed requests / wall_s */
    double peak_bytes;                      /* admitted footprint high-water mark */
} p4g_ServerStats;

int p4g_server_create(size_t hbm_budget_bytes, const char *options, p4g_Server *srv);
/* device footprint of one request for solver s (bytes) */
int p4g_server_request_bytes(p4g_Solver s, size_t *bytes);
int p4g_server_submit(p4g_Server srv, p4g_Solver s, const double *b, double *x, long *id);
/* state: 0 queued, 1 starting, 2 iterating, 3 done, -1 failed */
int p4g_server_poll(p4g_Server srv, long id, int *state);
/* blocks until the request is done; rep->solve_ms is its latency,
 * rep->setup_ms its queueing time, rep->inner_its its FP64 replacements */
int p4g_server_wait(p4g_Server srv, long id, p4g_Report *rep);
int p4g_server_stats(p4g_Server srv, p4g_ServerStats *st);
int p4g_server_destroy(p4g_Server *srv);

/* ------------------------------------------------ PETSc binary files */

/* Vectors and the operator in PETSc's binary format (what VecView, VecLoad
 * and MatView write and read through a binary PetscViewer: big-endian
 * int32 class id, sizes, then the data), to exchange right-hand sides,
 * solutions and the matrix with PETSc programs such as c/ch2/loadsolve.c
 * (VecLoad/MatLoad, loadsolve.c:48-70) or the dump of c/ch12/obstacle.c:
 * 137-146 (VecView of u, then of the lower bound, into one file).  A file
 * holds its objects one after another; append 0 starts a new file.
 * Error codes: 65 cannot open, 66 short read, 67 write failed, 79 not a
 * PETSc Vec where one was expected (PETSC_ERR_FILE_*).
 * The solver options -ksp_view_mat, -ksp_view_rhs and -ksp_view_solution
 * binary:<file> write the operator, the right-hand side and the solution
 * after every KSP solve, as KSPSolve's viewers do. */
int p4g_vec_view_binary(const char *path, long n, const double *x, int append);
/* the index-th object of the file (0-based), which must be a Vec; x may be
 * NULL to query its length */
int p4g_vec_load_binary(const char *path, int index, long *n, double *x);
/* the operator the Jacobian callback (or the cell coefficient) defines on
 * this DMDA, as an AIJ matrix: sorted column indices per row, the entries
 * the callback inserts (MatView of the assembled Jacobian); host only */
int p4g_da_mat_view_binary(p4g_DA da, const char *path, int append);
/* MatIsSymmetric(J, tol) of the matrix the Jacobian callback assembles on
 * this DMDA at the iterate u (host, the grid's size; NULL: zero): every
 * entry it inserts (the last value at a position, as INSERT_VALUES keeps)
 * against its transpose, |a_ij - a_ji| <= tol, before any compression into
 * the device stencil.  What -mat_is_symmetric tol reports after each
 * assembly (c/ch6/makefile:14,20,33; c/ch12/makefile:17); host only. */
int p4g_da_mat_is_symmetric(p4g_DA da, const double *u, double tol, int *symmetric);

/* ---------------------------------------------------- device utilities */

int   p4g_device_count(int *n);
int   p4g_device_synchronize(void);
void *p4g_device_malloc(size_t bytes);
int   p4g_device_free(void *p);
int   p4g_memcpy_htod(void *dst, const void *src, size_t bytes);
int   p4g_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int   p4g_memcpy_dtod(void *dst, const void *src, size_t bytes);   /* device-to-device copy */
const char *p4g_last_error(void);
const char *p4g_version(void);

#ifdef __cplusplus
}
#endif
#endif
