/*
 * minimal.h -- the minimal-surface equation of c/ch7/minimal.c on the p4g C
 * ABI: -div((1 + |grad u|^2)^q grad u) = 0 on the unit square, u = g
 * (tent or catenoid) on the boundary, Newton's method with colored finite-
 * difference Jacobians (p4g_ms_solve), optional grid sequencing and the
 * area monitor.  minimal_run() takes the reference's options and prints
 * what the reference prints.
 */
#ifndef P4G_MINIMAL_H
#define P4G_MINIMAL_H

#include "p4g.h"

#ifdef __cplusplus
extern "C" {
#endif

/* same layout as the oracle's minimal_result */
typedef struct {
    int    mx, my;
    int    its, reason, last_ksp_its, ksp_its_total, stages;
    int    stage_its[16];
    double fnorm0, fnorm, errinf, area;
    double setup_ms, solve_ms;
} minimal_result;

/* Run minimal with argv.  Text to outbuf (stdout if NULL); the final
 * solution (mx*my doubles) to u_out if not NULL.  Returns 0 or an error
 * code (PETSc numbering). */
int minimal_run(int argc, const char **argv, minimal_result *res, double *u_out, char *outbuf, int outlen);

#ifdef __cplusplus
}
#endif
#endif
