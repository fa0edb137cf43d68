/*
 * minimal.c -- the minimal-surface driver on the p4g C ABI, following
 * c/ch7/minimal.c:58-205: a 2D DMDA from 3x3 points (-da_grid_x/y,
 * -da_refine) on the unit square, the residual FormFunctionLocal with the
 * tent or catenoid boundary values (-ms_problem, -ms_q, -ms_tent_H,
 * -ms_catenoid_c), Newton's method from u = g on the boundary and 0 inside
 * (or the catenoid, -ms_exact_init), grid sequencing by Q1 interpolation
 * (-snes_grid_sequence), the area monitor (-ms_monitor, -ms_quaddegree) and
 * the report "done on ...".
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "minimal.h"
#include "poissonfunctions.h"

typedef struct {
    double q, tent_H, catenoid_c;
    int quaddegree;
} MinimalCtx;

/* boundary values (minimal.c:29-45) */
static PetscReal g_bdry_tent(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    const MinimalCtx *m = (const MinimalCtx *)((PoissonCtx *)ctx)->addctx;
    (void)z;
    return x < 1.0e-8 ? 2.0 * m->tent_H * (y < 0.5 ? y : 1.0 - y) : 0.0;
}

static PetscReal g_bdry_catenoid(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    const MinimalCtx *m = (const MinimalCtx *)((PoissonCtx *)ctx)->addctx;
    const double c = m->catenoid_c, ch = cosh(x / c);
    (void)z;
    return c * ch * sin(acos((y / c) / ch));
}

static double DD(double w, double q) { return pow(1.0 + w, q); }

/* the residual (minimal.c:230-301): Dirichlet rows u - g, interior rows the
 * flux differences with the face diffusivities (1 + |grad u|^2)^q, with
 * boundary neighbours taken from g */
static PetscErrorCode FormFunctionLocal(DMDALocalInfo *info, PetscReal **au, PetscReal **FF, PoissonCtx *user) {
    const MinimalCtx *m = (const MinimalCtx *)user->addctx;
    PetscReal lo[2], hi[2];
    DMGetBoundingBox(info->da, lo, hi);
    const double hx = (hi[0] - lo[0]) / (info->mx - 1), hy = (hi[1] - lo[1]) / (info->my - 1);
    const int mx = info->mx, my = info->my;
    for (int j = info->ys; j < info->ys + info->ym; j++) {
        const double y = j * hy;
        for (int i = info->xs; i < info->xs + info->xm; i++) {
            const double x = i * hx;
            if (j == 0 || i == 0 || i == mx - 1 || j == my - 1) {
                FF[j][i] = au[j][i] - user->g_bdry(x, y, 0.0, user);
                continue;
            }
            const int E = i + 1 == mx - 1, W = i - 1 == 0, N = j + 1 == my - 1, S = j - 1 == 0;
            const double uc = au[j][i];
            const double ue = E ? user->g_bdry(x + hx, y, 0.0, user) : au[j][i + 1];
            const double uw = W ? user->g_bdry(x - hx, y, 0.0, user) : au[j][i - 1];
            const double un = N ? user->g_bdry(x, y + hy, 0.0, user) : au[j + 1][i];
            const double us = S ? user->g_bdry(x, y - hy, 0.0, user) : au[j - 1][i];
            const double une = (E || N) ? user->g_bdry(x + hx, y + hy, 0.0, user) : au[j + 1][i + 1];
            const double unw = (W || N) ? user->g_bdry(x - hx, y + hy, 0.0, user) : au[j + 1][i - 1];
            const double use = (E || S) ? user->g_bdry(x + hx, y - hy, 0.0, user) : au[j - 1][i + 1];
            const double usw = (W || S) ? user->g_bdry(x - hx, y - hy, 0.0, user) : au[j - 1][i - 1];
            double gx, gy;
            gx = (ue - uc) / hx;
            gy = (un + une - us - use) / (4.0 * hy);
            const double De = DD(gx * gx + gy * gy, m->q);
            gx = (uc - uw) / hx;
            gy = (unw + un - usw - us) / (4.0 * hy);
            const double Dw = DD(gx * gx + gy * gy, m->q);
            gx = (ue + une - uw - unw) / (4.0 * hx);
            gy = (un - uc) / hy;
            const double Dn = DD(gx * gx + gy * gy, m->q);
            gx = (ue + use - uw - usw) / (4.0 * hx);
            gy = (uc - us) / hy;
            const double Ds = DD(gx * gx + gy * gy, m->q);
            FF[j][i] = -(hy / hx) * (De * (ue - uc) - Dw * (uc - uw)) - (hx / hy) * (Dn * (un - uc) - Ds * (uc - us));
        }
    }
    return 0;
}

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
static double opt_real(const Args *a, const char *k, double d) {
    const char *v = opt_get(a, k);
    return (v && *v) ? atof(v) : d;
}
static int opt_bool(const Args *a, const char *k) {
    const char *v = opt_get(a, k);
    return v && !(strcmp(v, "false") == 0 || strcmp(v, "0") == 0 || strcmp(v, "no") == 0);
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

static void petsc_g(char *b, int n, double v) {   /* PetscPrintf's %g keeps "0." */
    snprintf(b, n, "%g", v);
    if (!strpbrk(b, ".eEin")) strncat(b, ".", n - strlen(b) - 1);
}

static void snes_line(Out *o, int ind, int it, double f) {   /* SNESMonitorDefaultShort */
    if (f > 1e-9) outf(o, "%*s%3d SNES Function norm %g\n", ind, "", it, f);
    else if (f > 1e-11) outf(o, "%*s%3d SNES Function norm %5.3e\n", ind, "", it, f);
    else outf(o, "%*s%3d SNES Function norm < 1.e-11\n", ind, "", it);
}

static const char *ksp_reason_name(int r) {
    switch (r) {
    case P4G_CONVERGED_RTOL: return "CONVERGED_RTOL";
    case P4G_CONVERGED_ATOL: return "CONVERGED_ATOL";
    case P4G_CONVERGED_ITS: return "CONVERGED_ITS";
    case P4G_DIVERGED_NULL: return "DIVERGED_NULL";
    case P4G_DIVERGED_ITS: return "DIVERGED_ITS";
    case P4G_DIVERGED_DTOL: return "DIVERGED_DTOL";
    case P4G_DIVERGED_BREAKDOWN: return "DIVERGED_BREAKDOWN";
    case P4G_DIVERGED_INDEFINITE_PC: return "DIVERGED_INDEFINITE_PC";
    case P4G_DIVERGED_INDEFINITE_MAT: return "DIVERGED_INDEFINITE_MAT";
    case P4G_DIVERGED_NANORINF: return "DIVERGED_NANORINF";
    }
    return "CONVERGED_ITERATING";
}

static const char *snes_reason_name(int r) {
    switch (r) {
    case 2: return "CONVERGED_FNORM_ABS";
    case 3: return "CONVERGED_FNORM_RELATIVE";
    case 4: return "CONVERGED_SNORM_RELATIVE";
    case 5: return "CONVERGED_ITS";
    case -3: return "DIVERGED_LINEAR_SOLVE";
    case -4: return "DIVERGED_FNORM_NAN";
    case -5: return "DIVERGED_MAX_IT";
    case -6: return "DIVERGED_LINE_SEARCH";
    case -9: return "DIVERGED_DTOL";
    }
    return "CONVERGED_ITERATING";
}

/* MSEMonitor (minimal.c:305-362): area of the Q1 surface by tensor Gauss-
 * Legendre quadrature (c/interlude/quadrature.h's rules) and the extreme
 * diffusivities at the quadrature points */
static void mse_area(const double *u, int mx, int my, const MinimalCtx *m, double *area, double *Dmin,
                     double *Dmax) {
    static const double xi[3][3] = {{0.0, 0.0, 0.0}, {-0.577350269189626, 0.577350269189626, 0.0},
                                    {-0.774596669241483, 0.0, 0.774596669241483}};
    static const double wt[3][3] = {{2.0, 0.0, 0.0}, {1.0, 1.0, 0.0},
                                    {0.555555555555556, 0.888888888888889, 0.555555555555556}};
    const int r0 = m->quaddegree - 1, nq = m->quaddegree;
    const double hx = 1.0 / (mx - 1), hy = 1.0 / (my - 1);
    double a = 0.0, dmin = INFINITY, dmax = 0.0;
    for (int j = 1; j < my; j++) {
        const double yj = j * hy;
        for (int i = 1; i < mx; i++) {
            const double xi_ = i * hx;
            const double u11 = u[(long)j * mx + i], u01 = u[(long)j * mx + i - 1];
            const double u10 = u[(long)(j - 1) * mx + i], u00 = u[(long)(j - 1) * mx + i - 1];
            for (int r = 0; r < nq; r++) {
                const double x = xi_ - hx + hx * 0.5 * (xi[r0][r] + 1);
                for (int s = 0; s < nq; s++) {
                    const double y = yj - hy + hy * 0.5 * (xi[r0][s] + 1);
                    double ux = (u11 - u01) * (y - (yj - hy)) + (u10 - u00) * (yj - y);
                    ux /= hx * hy;
                    double uy = (u11 - u10) * (x - (xi_ - hx)) + (u01 - u00) * (xi_ - x);
                    uy /= hx * hy;
                    const double W = ux * ux + uy * uy, D = DD(W, m->q);
                    if (D < dmin) dmin = D;
                    if (D > dmax) dmax = D;
                    a += wt[r0][r] * wt[r0][s] * sqrt(1.0 + W);
                }
            }
        }
    }
    *area = a * (hx * hy / 4.0);
    *Dmin = dmin;
    *Dmax = dmax;
}

/* the area lines of one stage, kept until the history is printed */
typedef struct {
    const MinimalCtx *m;
    int mx, my, ind, n;
    char (*line)[96];
} MonCtx;

static int mse_monitor(void *ctx, int it, double fnorm, const double *x) {
    MonCtx *c = (MonCtx *)ctx;
    (void)fnorm;
    if (it >= c->n) return 0;
    double a, d0, d1;
    mse_area(x, c->mx, c->my, c->m, &a, &d0, &d1);
    snprintf(c->line[it], sizeof c->line[it], "%*sarea = %.8f; %.4f <= D <= %.4f\n", c->ind, "", a, d0, d1);
    return 0;
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
static double *interp_q1(const double *uc, int mxc, int mxf, int myf) {
    double *uf = (double *)malloc(sizeof(double) * (size_t)mxf * myf);
    for (int j = 0; j < myf; j++)
        for (int i = 0; i < mxf; i++) {
            const int i0 = i / 2, j0 = j / 2, i1 = i0 + (i & 1), j1 = j0 + (j & 1);
            uf[(long)j * mxf + i] = 0.25 * (uc[(long)j0 * mxc + i0] + uc[(long)j0 * mxc + i1] +
                                            uc[(long)j1 * mxc + i0] + uc[(long)j1 * mxc + i1]);
        }
    return uf;
}

int minimal_run(int argc, const char **argv, minimal_result *res, double *u_out, char *outbuf, int outlen) {
    Out o = {outbuf, outlen, 0};
    if (outbuf && outlen > 0) outbuf[0] = 0;
    minimal_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    Args a = {argc, argv};
    MinimalCtx mctx = {opt_real(&a, "ms_q", -0.5), opt_real(&a, "ms_tent_H", 1.0),
                       opt_real(&a, "ms_catenoid_c", 1.1), opt_int(&a, "ms_quaddegree", 3)};
    const int exact_init = opt_bool(&a, "ms_exact_init"), msmon = opt_bool(&a, "ms_monitor");
    const char *pb = opt_get(&a, "ms_problem");
    int tent = 0;
    if (pb) {
        if (!strcmp(pb, "tent")) tent = 1;
        else if (strcmp(pb, "catenoid")) {
            outf(&o, "[0]PETSC ERROR: unknown -ms_problem %s\n", pb);
            return 62;
        }
    }
    if (mctx.quaddegree < 1 || mctx.quaddegree > 3) {
        outf(&o, "[0]PETSC ERROR: -ms_quaddegree must be 1, 2 or 3\n");
        return PETSC_ERR_ARG_OUTOFRANGE;
    }
    if (tent && exact_init) {   /* minimal.c:107-129 */
        outf(&o, "[0]PETSC ERROR: initialization with exact solution only possible for -mse_problem catenoid\n");
        return 2;
    }
    if (!tent && mctx.catenoid_c < 1.0) {
        outf(&o, "[0]PETSC ERROR: catenoid exact solution only valid if c >= 1\n");
        return 3;
    }
    if (!tent && exact_init && mctx.q != -0.5) {
        outf(&o, "[0]PETSC ERROR: initialization with catenoid exact solution only possible if q=-0.5\n");
        return 4;
    }
    const int refine = opt_int(&a, "da_refine", 0), seq = opt_int(&a, "snes_grid_sequence", 0);
    if (seq < 0 || seq > 12) {
        outf(&o, "[0]PETSC ERROR: -snes_grid_sequence %d out of range\n", seq);
        return PETSC_ERR_ARG_OUTOFRANGE;
    }
    const int monitor = opt_get(&a, "snes_monitor_short") || opt_get(&a, "snes_monitor");
    const int kreason = opt_get(&a, "ksp_converged_reason") != NULL;
    const int sreason = opt_get(&a, "snes_converged_reason") != NULL;
    const char *sv = opt_get(&a, "mat_is_symmetric");
    char symtol[32] = "";
    if (sv) petsc_g(symtol, sizeof symtol, *sv ? atof(sv) : 0.0);
    PoissonCtx user;
    memset(&user, 0, sizeof user);
    user.Lx = user.Ly = user.Lz = 1.0;
    user.cx = user.cy = user.cz = 1.0;
    user.g_bdry = tent ? g_bdry_tent : g_bdry_catenoid;
    user.addctx = &mctx;
    double *u = NULL;
    int pmx = 0, e = 0;
    res->stages = seq + 1;
    for (int stage = 0; stage <= seq && !e; stage++) {
        char *opts = join_refine(argc, argv, refine + stage);
        DM da = NULL;
        p4g_Solver s = NULL;
        e = p4g_da_create(2, 3, 3, 1, 1.0, 1.0, 1.0, &da);
        if (!e) e = DMDASetUniformCoordinates(da, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
        if (!e) e = p4g_da_set_from_options(da, opts);
        if (!e) e = DMSetApplicationContext(da, &user);
        if (!e) e = p4g_da_set_function_local(da, (DMDASNESFunctionFn *)&FormFunctionLocal, &user);
        /* minimal.c:159-160 registers the Poisson Jacobian; -snes_fd_color replaces it */
        if (!e) e = p4g_da_set_jacobian_local(da, (DMDASNESJacobianFn *)&Poisson2DJacobianLocal, &user);
        if (!e) e = p4g_solver_create(da, opts, &s);
        free(opts);
        int mx = 0, my = 0;
        if (!e) e = p4g_da_get_sizes(da, NULL, &mx, &my, NULL, NULL);
        const long n = (long)mx * my;
        const double hx = mx > 1 ? 1.0 / (mx - 1) : 1.0, hy = my > 1 ? 1.0 / (my - 1) : 1.0;
        double *gb = NULL;
        char (*lines)[96] = NULL;
        if (!e) {
            gb = (double *)malloc(sizeof(double) * n);
            for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) gb[(long)j * mx + i] = user.g_bdry(i * hx, j * hy, 0.0, &user);
            if (stage == 0) {   /* InitialState(ZEROS, g on the boundary) or FormExactFromG */
                u = (double *)calloc(n, sizeof(double));
                for (int j = 0; j < my; j++)
                    for (int i = 0; i < mx; i++)
                        if (exact_init || i == 0 || j == 0 || i == mx - 1 || j == my - 1) u[(long)j * mx + i] = gb[(long)j * mx + i];
            } else {
                double *uf = interp_q1(u, pmx, mx, my);
                free(u);
                u = uf;
            }
            const int ind = 2 * (seq - stage);
            lines = (char (*)[96])calloc(64, sizeof *lines);
            MonCtx mc = {&mctx, mx, my, ind, 64, lines};
            p4g_MSReport rep;
            e = p4g_ms_solve(s, mctx.q, gb, u, msmon ? mse_monitor : NULL, &mc, &rep);
            if (!e) {
                int nh = 0, nk = 0;
                p4g_ms_get_history(s, NULL, NULL, NULL, NULL, 0, &nh, &nk);
                const int cap = (nh > nk ? nh : nk) + 1;
                double *fh = (double *)malloc(sizeof(double) * cap);
                int *ki = (int *)malloc(sizeof(int) * cap), *kr = (int *)malloc(sizeof(int) * cap);
                int *js = (int *)malloc(sizeof(int) * cap);
                p4g_ms_get_history(s, fh, ki, kr, js, cap, &nh, &nk);
                if (sv) outf(&o, "Matrix is symmetric (tolerance %s)\n", symtol);   /* DMCreateMatrix's assembly */
                for (int it = 0; it < nh || it <= nk; it++) {
                    if (it > 0) {   /* Newton step it-1: its Jacobian and linear solve */
                        if (it - 1 >= nk) break;
                        if (sv) outf(&o, "Matrix is %ssymmetric (tolerance %s)\n", js[it - 1] == 0 ? "not " : "", symtol);
                        if (kreason) {
                            if (kr[it - 1] > 0)
                                outf(&o, "%*s    Linear solve converged due to %s iterations %d\n", ind, "",
                                     ksp_reason_name(kr[it - 1]), ki[it - 1]);
                            else
                                outf(&o, "%*s    Linear solve did not converge due to %s iterations %d\n", ind, "",
                                     ksp_reason_name(kr[it - 1]), ki[it - 1]);
                        }
                    }
                    if (it >= nh) break;
                    if (msmon && it < 64) outf(&o, "%s", lines[it]);
                    if (monitor) snes_line(&o, ind, it, fh[it]);
                }
                free(fh); free(ki); free(kr); free(js);
                if (stage < 16) res->stage_its[stage] = rep.its;
                res->ksp_its_total += rep.ksp_its_total;
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
                if (stage == seq) {   /* report (minimal.c:184-204) */
                    double d0, d1;
                    res->mx = mx; res->my = my;
                    res->its = rep.its; res->reason = rep.reason; res->last_ksp_its = rep.last_ksp_its;
                    res->fnorm0 = rep.fnorm0; res->fnorm = rep.fnorm;
                    mse_area(u, mx, my, &mctx, &res->area, &d0, &d1);
                    outf(&o, "done on %d x %d grid and problem %s", mx, my, tent ? "tent" : "catenoid");
                    if (!tent && mctx.q == -0.5) {
                        double emax = 0.0;
                        for (long k = 0; k < n; k++) {
                            const double d = fabs(u[k] - gb[k]);   /* FormExactFromG: g at every node */
                            if (d > emax) emax = d;
                        }
                        res->errinf = emax;
                        outf(&o, ":  error |u-uexact|_inf = %.5e\n", emax);
                    } else {
                        outf(&o, " ...\n");
                    }
                    if (u_out) memcpy(u_out, u, sizeof(double) * n);
                }
            }
        }
        if (e) outf(&o, "[0]PETSC ERROR: %s\n", p4g_last_error());
        pmx = mx;
        free(gb);
        free(lines);
        p4g_solver_destroy(&s);
        p4g_da_destroy(&da);
    }
    free(u);
    return e;
}
