/*
 * obstacle.h -- the obstacle problem of c/ch12/obstacle.c on the p4g C ABI:
 * -Lap u = 0, u >= psi on (-2,2)^2, u = u_exact on the boundary, solved by
 * the reduced-space VI Newton method (p4g_vi_solve) over the ch6 Poisson
 * callbacks, with optional grid sequencing.  obstacle_run() takes the
 * reference's options and prints what the reference prints.
 */
#ifndef P4G_OBSTACLE_H
#define P4G_OBSTACLE_H

#include "p4g.h"

#ifdef __cplusplus
extern "C" {
#endif

/* same layout as the oracle's obstacle_result */
typedef struct {
    int    mx, my;
    int    its, reason, last_ksp_its, ksp_its_total, stages;
    int    stage_its[16];
    long   active;
    double fnorm0, fnorm, error1, errorinf, areaerr;
    double setup_ms, solve_ms;
} obstacle_result;

/* Run obstacle with argv.  Text to outbuf (stdout if NULL); the final
 * solution (mx*my doubles) to u_out if not NULL.  Returns 0 or an error
 * code (PETSc numbering). */
int obstacle_run(int argc, const char **argv, obstacle_result *res, double *u_out, char *outbuf,
                 int outlen);

#ifdef __cplusplus
}
#endif
#endif
