/*
 * obstacle.c -- the obstacle problem driver on the p4g C ABI, following
 * c/ch12/obstacle.c:71-199: a 2D DMDA from 3x3 points (-da_grid_x/y,
 * -da_refine) on (-2,2)^2 (DMDASetUniformCoordinates), the ch6
 * Poisson2DFunctionLocal/JacobianLocal callbacks with g = u_exact and
 * f = 0, bounds psi <= u (FormBounds), SNESVINEWTONRSLS from u = 0, grid
 * sequencing by Q1 interpolation (-snes_grid_sequence), and the report
 * "done on ..." plus the errors and the active-area error (GetActiveSet).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "obstacle.h"
#include "poissonfunctions.h"

/* ------------------------------------------------ problem (obstacle.c:15-69) */

static double psi(double x, double y) {
    const double r = sqrt(x * x + y * y), r0 = 0.9, psi0 = sqrt(1.0 - r0 * r0), dpsi0 = -r0 / psi0;
    return r <= r0 ? sqrt(1.0 - r * r) : psi0 + dpsi0 * (r - r0);
}

/* radial free-boundary solution, free boundary at r = afree */
static double u_exact(double x, double y) {
    const double afree = 0.697965148223374, A = 0.680259411891719, B = 0.471519893402112;
    const double r = sqrt(x * x + y * y);
    return r <= afree ? psi(x, y) : -A * log(r) + B;
}

static PetscReal g_fcn(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    (void)z; (void)ctx;
    return u_exact(x, y);
}

static PetscReal zero(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    (void)x; (void)y; (void)z; (void)ctx;
    return 0.0;
}

#define AEXACT 0.697965148223374

/* ---------------------------------------------------------------- options */

typedef struct { int argc; const char **argv; } Args;

static int is_val(const char *t) { return t[0] != '-' || (t[1] >= '0' && t[1] <= '9') || t[1] == '.'; }

static const char *opt_get(const Args *a, const char *key) {   /* key without '-' */
    const char *hit = NULL;
    for (int q = 0; q < a->argc; q++) {
        const char *t = a->argv[q];
        if (t[0] != '-' || is_val(t) || strcmp(t + 1, key)) continue;
        hit = (q + 1 < a->argc && is_val(a->argv[q + 1])) ? a->argv[q + 1] : "";
    }
    return hit;
}
static int opt_int(const Args *a, const char *k, int d) {
    const char *v = opt_get(a, k);
    return (v && *v) ? atoi(v) : d;
}

typedef struct { char *buf; int len, pos; } Out;
static void outf(Out *o, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (o->buf) {
        if (o->pos < o->len) {
            int w = vsnprintf(o->buf + o->pos, (size_t)(o->len - o->pos), fmt, ap);
            if (w > 0) o->pos += (w < o->len - o->pos) ? w : o->len - o->pos - 1;
        }
    } else {
        vprintf(fmt, ap);
    }
    va_end(ap);
}

/* PetscPrintf's %g keeps a '.' on an integral value ("0.") */
static void petsc_g(char *b, int n, double v) {
    snprintf(b, n, "%g", v);
    if (!strpbrk(b, ".eEin")) strncat(b, ".", n - strlen(b) - 1);
}

/* SNESMonitorDefaultShort */
static void snes_line(Out *o, int ind, int it, double f) {
    if (f > 1e-9) outf(o, "%*s%3d SNES Function norm %g\n", ind, "", it, f);
    else if (f > 1e-11) outf(o, "%*s%3d SNES Function norm %5.3e\n", ind, "", it, f);
    else outf(o, "%*s%3d SNES Function norm < 1.e-11\n", ind, "", it);
}

static const char *ksp_reason_name(int r) {
    switch (r) {
    case P4G_CONVERGED_RTOL: return "CONVERGED_RTOL";
    case P4G_CONVERGED_ATOL: return "CONVERGED_ATOL";
    case P4G_CONVERGED_ITS: return "CONVERGED_ITS";
    case P4G_DIVERGED_ITS: return "DIVERGED_ITS";
    case P4G_DIVERGED_DTOL: return "DIVERGED_DTOL";
    case P4G_DIVERGED_BREAKDOWN: return "DIVERGED_BREAKDOWN";
    case P4G_DIVERGED_INDEFINITE_PC: return "DIVERGED_INDEFINITE_PC";
    case P4G_DIVERGED_INDEFINITE_MAT: return "DIVERGED_INDEFINITE_MAT";
    case P4G_DIVERGED_NANORINF: return "DIVERGED_NANORINF";
    }
    return "CONVERGED_ITERATING";
}

/* SNESConvergedReasons[] */
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

static char *join_refine(int argc, const char **argv, int refine) {
    size_t n = 64;
    for (int q = 0; q < argc; q++) n += strlen(argv[q]) + 1;
    char *s = (char *)malloc(n);
    s[0] = 0;
    for (int q = 0; q < argc; q++) { strcat(s, argv[q]); strcat(s, " "); }
    char tail[48];
    snprintf(tail, sizeof tail, "-da_refine %d", refine);   /* the last -da_refine wins */
    strcat(s, tail);
    return s;
}

/* MatInterpolate with the Q1 interpolation of a vertex-nested 2D grid */
static double *interp_q1(const double *uc, int mxc, int myc, int mxf, int myf) {
    double *uf = (double *)malloc(sizeof(double) * (size_t)mxf * myf);
    for (int j = 0; j < myf; j++)
        for (int i = 0; i < mxf; i++) {
            const int i0 = i / 2, j0 = j / 2, i1 = i0 + (i & 1), j1 = j0 + (j & 1);
            uf[(long)j * mxf + i] = 0.25 * (uc[(long)j0 * mxc + i0] + uc[(long)j0 * mxc + i1] +
                                            uc[(long)j1 * mxc + i0] + uc[(long)j1 * mxc + i1]);
        }
    (void)myc;
    return uf;
}

int obstacle_run(int argc, const char **argv, obstacle_result *res, double *u_out, char *outbuf, int outlen) {
    Out o = {outbuf, outlen, 0};
    if (outbuf && outlen > 0) outbuf[0] = 0;
    obstacle_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    Args a = {argc, argv};
    const int refine = opt_int(&a, "da_refine", 0), seq = opt_int(&a, "snes_grid_sequence", 0);
    if (seq < 0 || seq > 12) {
        outf(&o, "[0]PETSC ERROR: -snes_grid_sequence %d out of range\n", seq);
        return PETSC_ERR_ARG_OUTOFRANGE;
    }
    const int monitor = opt_get(&a, "snes_monitor_short") || opt_get(&a, "snes_monitor");
    const int kreason = opt_get(&a, "ksp_converged_reason") != NULL;
    const int sreason = opt_get(&a, "snes_converged_reason") != NULL;
    const char *sv = opt_get(&a, "mat_is_symmetric");
    const char *dump = opt_get(&a, "obs_dump_binary");   /* default name as obstacle.c:77 */
    if (dump && !*dump) dump = "dump.dat";
    char symtol[32] = "";
    if (sv) petsc_g(symtol, sizeof symtol, *sv ? atof(sv) : 0.0);
    PoissonCtx user;   /* obstacle.c:97-102 */
    memset(&user, 0, sizeof user);
    user.Lx = user.Ly = 4.0; user.Lz = 1.0;
    user.cx = user.cy = user.cz = 1.0;
    user.g_bdry = g_fcn;
    user.f_rhs = zero;
    double *u = NULL;
    int pmx = 0, pmy = 0, e = 0;
    res->stages = seq + 1;
    for (int stage = 0; stage <= seq && !e; stage++) {
        char *opts = join_refine(argc, argv, refine + stage);
        DM da = NULL;
        p4g_Solver s = NULL;
        e = p4g_da_create(2, 3, 3, 1, 4.0, 4.0, 1.0, &da);
        if (!e) e = DMDASetUniformCoordinates(da, -2.0, 2.0, -2.0, 2.0, -1.0, -1.0);
        if (!e) e = p4g_da_set_from_options(da, opts);
        if (!e) e = DMSetApplicationContext(da, &user);
        if (!e) e = p4g_da_set_function_local(da, (DMDASNESFunctionFn *)&Poisson2DFunctionLocal, &user);
        if (!e) e = p4g_da_set_jacobian_local(da, (DMDASNESJacobianFn *)&Poisson2DJacobianLocal, &user);
        if (!e) e = p4g_solver_create(da, opts, &s);
        free(opts);
        int mx = 0, my = 0;
        if (!e) e = p4g_da_get_sizes(da, NULL, &mx, &my, NULL, NULL);
        const long n = (long)mx * my;
        double *xl = NULL, *F = NULL;
        if (!e) {
            const double hx = 4.0 / (mx - 1), hy = 4.0 / (my - 1);
            xl = (double *)malloc(sizeof(double) * n);
            F = (double *)malloc(sizeof(double) * n);
            for (int j = 0; j < my; j++)   /* FormBounds (obstacle.c:234-253) */
                for (int i = 0; i < mx; i++) xl[(long)j * mx + i] = psi(-2.0 + i * hx, -2.0 + j * hy);
            if (stage == 0) {
                u = (double *)calloc(n, sizeof(double));   /* obstacle.c:130-131 */
            } else {
                double *uf = interp_q1(u, pmx, pmy, mx, my);
                free(u);
                u = uf;
            }
            p4g_VIReport rep;
            /* -mat_is_symmetric: the Jacobian the callback assembles, checked at
             * the stage's first iterate (Poisson2DJacobianLocal does not depend
             * on the iterate, so every Newton step assembles the same matrix) */
            int jac_sym = 1;
            if (sv) e = p4g_da_mat_is_symmetric(da, u, *sv ? atof(sv) : 0.0, &jac_sym);
            if (!e) e = p4g_vi_solve(s, xl, NULL, u, F, &rep);
            const int ind = 2 * (seq - stage);
            if (!e) {
                int nh = 0, nk = 0;
                p4g_vi_get_history(s, NULL, NULL, NULL, 0, &nh, &nk);
                const int cap = (nh > nk ? nh : nk) + 1;
                double *fh = (double *)malloc(sizeof(double) * cap);
                int *ki = (int *)malloc(sizeof(int) * cap), *kr = (int *)malloc(sizeof(int) * cap);
                p4g_vi_get_history(s, fh, ki, kr, cap, &nh, &nk);
                if (sv) outf(&o, "%*sMatrix is symmetric (tolerance %s)\n", ind, "", symtol);   /* preallocated, zero */
                if (monitor && nh > 0) snes_line(&o, ind, 0, fh[0]);
                for (int q = 0; q < nk; q++) {
                    if (sv) outf(&o, "%*sMatrix is %ssymmetric (tolerance %s)\n", ind, "", jac_sym ? "" : "not ",
                                 symtol);
                    if (kreason) {
                        if (kr[q] > 0)
                            outf(&o, "%*s    Linear solve converged due to %s iterations %d\n", ind, "",
                                 ksp_reason_name(kr[q]), ki[q]);
                        else
                            outf(&o, "%*s    Linear solve did not converge due to %s iterations %d\n", ind, "",
                                 ksp_reason_name(kr[q]), ki[q]);
                    }
                    if (monitor && q + 1 < nh) snes_line(&o, ind, q + 1, fh[q + 1]);
                }
                free(fh); free(ki); free(kr);
                if (stage < 16) res->stage_its[stage] = rep.its;
                if (sreason) {
                    if (rep.reason > 0)
                        outf(&o, "%*s  Nonlinear solve converged due to %s iterations %d\n", ind, "",
                             snes_reason_name(rep.reason), rep.its);
                    else
                        outf(&o, "%*s  Nonlinear solve did not converge due to %s iterations %d\n", ind, "",
                             snes_reason_name(rep.reason), rep.its);
                }
                res->setup_ms += rep.setup_ms;
                res->solve_ms += rep.solve_ms;
                if (stage == seq && dump) {   /* -obs_dump_binary: VecView of u, then of psi (obstacle.c:137-146) */
                    outf(&o, "writing u,psi in binary format to %s ...\n", dump);
                    e = p4g_vec_view_binary(dump, n, u, 0);
                    if (!e) e = p4g_vec_view_binary(dump, n, xl, 1);
                }
                if (stage == seq && !e) {   /* report (obstacle.c:160-198) */
                    res->mx = mx; res->my = my;
                    res->its = rep.its; res->reason = rep.reason; res->last_ksp_its = rep.last_ksp_its;
                    res->ksp_its_total = rep.ksp_its_total; res->fnorm0 = rep.fnorm0; res->fnorm = rep.fnorm;
                    if (rep.reason <= 0)
                        outf(&o, "WARNING: SNES not converged ... use -snes_converged_reason to check\n");
                    outf(&o, "done on %d x %d grid ... %s, SNES iters = %d, last KSP iters = %d\n", mx, my,
                         snes_reason_name(rep.reason), rep.its, rep.last_ksp_its);
                    const double dx = 4.0 / (mx - 1), dy = 4.0 / (my - 1);
                    long act = 0;
                    double e1 = 0.0, einf = 0.0;
                    for (int j = 0; j < my; j++)
                        for (int i = 0; i < mx; i++) {
                            const long q = (long)j * mx + i;
                            if (u[q] <= xl[q] + 1.0e-8 && F[q] > 0.0) act++;   /* GetActiveSet */
                            const double d = fabs(u[q] - u_exact(-2.0 + i * dx, -2.0 + j * dy));
                            e1 += d;
                            if (d > einf) einf = d;
                        }
                    const double actarea = dx * dy * act, exactarea = M_PI * AEXACT * AEXACT;
                    res->active = act;
                    res->error1 = e1 / ((double)mx * my);
                    res->errorinf = einf;
                    res->areaerr = fabs(actarea - exactarea) / exactarea;
                    outf(&o, "errors: av |u-uexact| = %.3e, |u-uexact|_inf = %.3e, active area error = %.3f%%\n",
                         res->error1, res->errorinf, 100.0 * res->areaerr);
                    if (u_out) memcpy(u_out, u, sizeof(double) * n);
                }
            }
        }
        if (e) outf(&o, "[0]PETSC ERROR: %s\n", p4g_last_error());
        pmx = mx; pmy = my;
        free(xl); free(F);
        p4g_solver_destroy(&s);
        p4g_da_destroy(&da);
    }
    free(u);
    return e;
}
