/*
 * fish.h -- the structured-grid Poisson driver of c/ch6/fish.c on the p4g
 * C ABI.  fish_run() takes the reference's command-line options (-fsh_*,
 * -da_*, -ksp_*, -pc_*, -mg_*) plus the p4g options and prints what the
 * reference prints.  A session keeps one problem resident on the GPU for
 * repeated timed solves.
 */
#ifndef P4G_FISH_H
#define P4G_FISH_H

#include "p4g.h"

#ifdef __cplusplus
extern "C" {
#endif

/* same layout as the oracle's fish_result */
typedef struct {
    int    dim, mx, my, mz;
    int    its, reason, inner_its, nlevels;
    double fnorm0, fnorm1, rnorm0, rnorm, true_relres, errinf, err2h;
    double setup_ms, solve_ms, emax_fine;
} fish_result;

/* Run fish with argv.  Output text goes to outbuf (stdout if NULL); the
 * solution (mx*my*mz doubles) to u_out if not NULL.  Returns 0 or the
 * reference's error code. */
int fish_run(int argc, const char **argv, fish_result *res, double *u_out, char *outbuf,
             int outlen);

typedef struct fish_session_s *fish_session;
/* problem, grid, solver, set-up and F(u0) resident on the device */
int fish_session_create(int argc, const char **argv, fish_session *s);
/* one KSP solve J y = F(u0) from y = 0 on device-resident data */
int fish_session_solve(fish_session s, p4g_Report *rep);
/* errors of u = u0 - y of the last solve against the exact solution */
int fish_session_errors(fish_session s, double *errinf, double *err2h);
/* copy u = u0 - y of the last solve to host memory (n doubles) */
int fish_session_solution(fish_session s, double *u);
p4g_Solver fish_session_solver(fish_session s);
int  fish_session_sizes(fish_session s, int *dim, int *mx, int *my, int *mz);
const double *fish_session_rhs_device(fish_session s);
int  fish_session_destroy(fish_session *s);
/* host only: set up the problem of the command line and write its operator
 * (the finest level, the rows the Jacobian callback inserts) to path in
 * PETSc's binary AIJ format, as -mat_view binary:<file> does */
int fish_write_operator(int argc, const char **argv, const char *path);

#ifdef __cplusplus
}
#endif
#endif
