/*
 * fish.c -- structured-grid Poisson driver on the p4g C ABI, following
 * c/ch6/fish.c:128-286: options with prefix fsh_, a DMDA grown from 3 points
 * per dimension by -da_refine, residual/Jacobian callbacks registered on the
 * grid, SNES KSPONLY with CG, InitialState, the solve, and the error report
 * |u-uexact|_inf and |u-uexact|_h = |u-uexact|_2 / sqrt(prod(m-1)).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fish.h"
#include "poissonfunctions.h"
#include "problems.h"

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
static double opt_real(const Args *a, const char *k, double d) {
    const char *v = opt_get(a, k);
    return (v && *v) ? atof(v) : d;
}
static int opt_int(const Args *a, const char *k, int d) {
    const char *v = opt_get(a, k);
    return (v && *v) ? atoi(v) : d;
}
static int opt_bool(const Args *a, const char *k, int d) {
    const char *v = opt_get(a, k);
    if (!v) return d;
    if (!*v) return 1;
    return !(strcmp(v, "false") == 0 || strcmp(v, "0") == 0 || strcmp(v, "no") == 0);
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

/* ------------------------------------------------------- problem context */

typedef struct {
    PoissonCtx  pc;      /* first: the callbacks receive &ctx->pc */
    FishProblem pr;
    VarCoefCtx  vc;
    double     *kcell;
} FishCtx;

static PetscReal g_cb(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    return fish_g(&((FishCtx *)ctx)->pr, x, y, z);
}
static PetscReal f_cb(PetscReal x, PetscReal y, PetscReal z, void *ctx) {
    return fish_f(&((FishCtx *)ctx)->pr, x, y, z);
}

/* residual_ptr / jacobian_ptr of fish.c:92-100 */
static DMDASNESFunctionFn *residual_ptr[3] = {(DMDASNESFunctionFn *)&Poisson1DFunctionLocal,
                                              (DMDASNESFunctionFn *)&Poisson2DFunctionLocal,
                                              (DMDASNESFunctionFn *)&Poisson3DFunctionLocal};
static DMDASNESJacobianFn *jacobian_ptr[3] = {(DMDASNESJacobianFn *)&Poisson1DJacobianLocal,
                                              (DMDASNESJacobianFn *)&Poisson2DJacobianLocal,
                                              (DMDASNESJacobianFn *)&Poisson3DJacobianLocal};

typedef struct {
    FishCtx    ctx;
    DM         da;
    p4g_Solver solver;
    int        initial_random, gonboundary;
    char      *optstr;
} Setup;

static void setup_free(Setup *s) {
    p4g_solver_destroy(&s->solver);
    p4g_da_destroy(&s->da);
    free(s->ctx.kcell);
    free(s->optstr);
    memset(s, 0, sizeof *s);
}

static char *join(int argc, const char **argv) {
    size_t n = 1;
    for (int q = 0; q < argc; q++) n += strlen(argv[q]) + 1;
    char *s = (char *)malloc(n);
    s[0] = 0;
    for (int q = 0; q < argc; q++) { strcat(s, argv[q]); strcat(s, " "); }
    return s;
}

static int setup(int argc, const char **argv, Setup *S, char *err, int errlen, int want_solver) {
    memset(S, 0, sizeof *S);
    Args a = {argc, argv};
    FishCtx *c = &S->ctx;
    FishProblem *p = &c->pr;
    /* fish defaults (fish.c:140-153) */
    p->dim = opt_int(&a, "fsh_dim", 2);
    p->problem = PR_MANUEXP;
    const char *pn = opt_get(&a, "fsh_problem");
    if (pn) {
        p->problem = -1;
        for (int q = 0; q < PR_COUNT; q++)
            if (!strcmp(pn, fish_problem_names[q])) p->problem = q;
        if (p->problem < 0) { snprintf(err, errlen, "unknown -fsh_problem %s\n", pn); return PETSC_ERR_ARG_WRONG; }
    }
    c->pc.Lx = opt_real(&a, "fsh_Lx", 1.0);
    c->pc.Ly = opt_real(&a, "fsh_Ly", 1.0);
    c->pc.Lz = opt_real(&a, "fsh_Lz", 1.0);
    c->pc.cx = p->cx = opt_real(&a, "fsh_cx", 1.0);
    c->pc.cy = p->cy = opt_real(&a, "fsh_cy", 1.0);
    c->pc.cz = p->cz = opt_real(&a, "fsh_cz", 1.0);
    const char *it = opt_get(&a, "fsh_initial_type");
    if (it && strcmp(it, "zeros") && strcmp(it, "random")) {
        snprintf(err, errlen, "unknown -fsh_initial_type %s\n", it);
        return PETSC_ERR_ARG_WRONG;
    }
    S->initial_random = it && !strcmp(it, "random");
    S->gonboundary = opt_bool(&a, "fsh_initial_gonboundary", 1);
    p->noise_sigma = opt_real(&a, "fsh_noise_sigma", 0.0);
    p->noise_seed = (uint64_t)opt_int(&a, "fsh_noise_seed", 0);
    p->modes_seed = (uint64_t)opt_int(&a, "fsh_modes_seed", 0);
    p->k_seed = (uint64_t)opt_int(&a, "fsh_k_seed", 0);
    /* argument checks in fish.c's order (fish.c:188-193, 214-215) */
    if (p->cx <= 0.0 || p->cy <= 0.0 || p->cz <= 0.0) {
        snprintf(err, errlen, "positivity required for coefficients cx,cy,cz\n");
        return 2;
    }
    if (p->problem == PR_MANUEXP && (p->cx != 1.0 || p->cy != 1.0 || p->cz != 1.0)) {
        snprintf(err, errlen, "cx=cy=cz=1 required for problem MANUEXP\n");
        return 3;
    }
    if (p->dim < 1 || p->dim > 3) {
        snprintf(err, errlen, "invalid dim for DMDA creation\n");
        return 1;
    }
    if (p->problem == PR_VARCOEF && p->dim != 3) {
        snprintf(err, errlen, "problem varcoef requires -fsh_dim 3\n");
        return PETSC_ERR_ARG_WRONG;
    }
    fish_problem_init(p);
    c->pc.g_bdry = g_cb;
    c->pc.f_rhs = f_cb;
    c->pc.addctx = NULL;
    /* DMDA from a 3-point base (fish.c:197-220) */
    S->optstr = join(argc, argv);
    int e = p4g_da_create(p->dim, 3, 3, 3, c->pc.Lx, c->pc.Ly, c->pc.Lz, &S->da);
    if (!e) e = p4g_da_set_from_options(S->da, S->optstr);
    if (e) { snprintf(err, errlen, "%s\n", p4g_last_error()); return e; }
    DMSetApplicationContext(S->da, &c->pc);
    int mx, my, mz;
    p4g_da_get_sizes(S->da, NULL, &mx, &my, &mz, NULL);
    p->mx = mx; p->my = my; p->mz = mz;
    p->h[0] = c->pc.Lx / (mx - 1);
    p->h[1] = p->dim > 1 ? c->pc.Ly / (my - 1) : 1.0;
    p->h[2] = p->dim > 2 ? c->pc.Lz / (mz - 1) : 1.0;
    if (p->problem == PR_VARCOEF) {
        const long nc = (long)(mx - 1) * (my - 1) * (mz - 1);
        c->kcell = (double *)malloc(sizeof(double) * nc);
        for (int k = 0; k < mz - 1; k++)
            for (int j = 0; j < my - 1; j++)
                for (int i = 0; i < mx - 1; i++)
                    c->kcell[((long)k * (my - 1) + j) * (mx - 1) + i] =
                        fish_k(p, (i + 0.5) * p->h[0], (j + 0.5) * p->h[1], (k + 0.5) * p->h[2]);
        c->vc.kcell = c->kcell;
        c->vc.mx = mx; c->vc.my = my; c->vc.mz = mz;
        c->pc.addctx = &c->vc;
        e = p4g_da_set_function_local(S->da, (DMDASNESFunctionFn *)&VarPoisson3DFunctionLocal, &c->pc);
        if (!e) e = p4g_da_set_cell_coefficient(S->da, c->kcell);
    } else {
        e = p4g_da_set_function_local(S->da, residual_ptr[p->dim - 1], &c->pc);
        if (!e) e = p4g_da_set_jacobian_local(S->da, jacobian_ptr[p->dim - 1], &c->pc);
    }
    if (!e && want_solver) e = p4g_solver_create(S->da, S->optstr, &S->solver);
    if (e) { snprintf(err, errlen, "%s\n", p4g_last_error()); return e; }
    return 0;
}

static void snes_line(Out *o, int it, double f) {
    if (f > 1e-9) outf(o, "%3d SNES Function norm %g\n", it, f);
    else if (f > 1e-11) outf(o, "%3d SNES Function norm %5.3e\n", it, f);
    else outf(o, "%3d SNES Function norm < 1.e-11\n", it);
}

static const char *reason_name(int r) {
    switch (r) {
    case P4G_CONVERGED_RTOL: return "CONVERGED_RTOL";
    case P4G_CONVERGED_ATOL: return "CONVERGED_ATOL";
    case P4G_CONVERGED_ITS: return "CONVERGED_ITS";
    case P4G_DIVERGED_ITS: return "DIVERGED_ITS";
    case P4G_DIVERGED_DTOL: return "DIVERGED_DTOL";
    case P4G_DIVERGED_NULL: return "DIVERGED_NULL";
    case P4G_DIVERGED_BREAKDOWN: return "DIVERGED_BREAKDOWN";
    case P4G_DIVERGED_INDEFINITE_PC: return "DIVERGED_INDEFINITE_PC";
    case P4G_DIVERGED_INDEFINITE_MAT: return "DIVERGED_INDEFINITE_MAT";
    case P4G_DIVERGED_NANORINF: return "DIVERGED_NANORINF";
    }
    return "CONVERGED_ITERATING";
}

/* |u - u_exact|_inf and |u - u_exact|_h (fish.c:248-276) */
static void errors(const Setup *S, const double *u, double *einf, double *e2h) {
    const FishProblem *p = &S->ctx.pr;
    const int mx = p->mx, my = p->dim > 1 ? p->my : 1, mz = p->dim > 2 ? p->mz : 1;
    double mi = 0.0, s2 = 0.0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                const double d = u[((long)k * my + j) * mx + i] -
                                 fish_g(p, i * p->h[0], p->dim > 1 ? j * p->h[1] : 0.0, p->dim > 2 ? k * p->h[2] : 0.0);
                mi = fabs(d) > mi ? fabs(d) : mi;
                s2 += d * d;
            }
    double nc = (double)(mx - 1);
    if (p->dim > 1) nc *= (my - 1);
    if (p->dim > 2) nc *= (mz - 1);
    *einf = mi;
    *e2h = sqrt(s2) / sqrt(nc);
}

int fish_run(int argc, const char **argv, fish_result *res, double *u_out, char *outbuf, int outlen) {
    Out o = {outbuf, outlen, 0};
    if (outbuf && outlen > 0) outbuf[0] = 0;
    fish_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    Setup S;
    char err[1200];
    int e = setup(argc, argv, &S, err, sizeof err, 1);
    if (e) {
        outf(&o, "[0]PETSC ERROR: %s", err);
        setup_free(&S);
        return e;
    }
    Args a = {argc, argv};
    /* initial iterate (fish.c:237-238) and SNESSolve (fish.c:239) */
    Vec u;
    DMGetGlobalVector(S.da, &u);
    e = InitialState(S.da, S.initial_random ? RANDOM : ZEROS, S.gonboundary ? PETSC_TRUE : PETSC_FALSE,
                     u, &S.ctx.pc);
    PetscScalar *ua = NULL;
    VecGetArray(u, &ua);
    const char *sym = opt_get(&a, "mat_is_symmetric");
    int jac_sym = 1;
    if (sym && !e) {   /* MatIsSymmetric of the Jacobian the callback assembles at u0 */
        e = p4g_da_mat_is_symmetric(S.da, ua, *sym ? atof(sym) : 0.0, &jac_sym);
    }
    p4g_Report rep;
    memset(&rep, 0, sizeof rep);
    if (!e) e = p4g_snes_solve(S.solver, ua, &rep);
    if (e) {
        outf(&o, "[0]PETSC ERROR: %s\n", p4g_last_error());
        DMRestoreGlobalVector(S.da, &u);
        setup_free(&S);
        return e;
    }
    const int monitor = opt_get(&a, "snes_monitor_short") || opt_get(&a, "snes_monitor");
    if (monitor) snes_line(&o, 0, rep.fnorm0);
    if (sym) {   /* PETSc reports after each assembly: the preallocated (zero, so
                    symmetric) matrix at SNES set-up, then the Jacobian */
        const double tol = *sym ? atof(sym) : 0.0;
        outf(&o, "Matrix is symmetric (tolerance %g)\n", tol);
        outf(&o, "Matrix is %ssymmetric (tolerance %g)\n", jac_sym ? "" : "not ", tol);
    }
    if (opt_get(&a, "ksp_monitor")) {   /* KSPMonitorResidual, at the KSP's tab level under SNES */
        int nh = 0;
        p4g_solver_get_history(S.solver, NULL, 0, &nh);
        double *h = (double *)malloc(sizeof(double) * (nh > 0 ? nh : 1));
        p4g_solver_get_history(S.solver, h, nh, &nh);
        for (int q = 0; q < nh; q++) outf(&o, "    %3d KSP Residual norm %14.12e \n", q, h[q]);
        free(h);
    }
    if (opt_get(&a, "ksp_converged_reason")) {
        if (rep.reason > 0)
            outf(&o, "    Linear solve converged due to %s iterations %d\n", reason_name(rep.reason), rep.its);
        else
            outf(&o, "    Linear solve did not converge due to %s iterations %d\n", reason_name(rep.reason), rep.its);
    }
    if (monitor) snes_line(&o, 1, rep.fnorm1);
    const FishProblem *p = &S.ctx.pr;
    res->dim = p->dim; res->mx = p->mx; res->my = p->dim > 1 ? p->my : 1; res->mz = p->dim > 2 ? p->mz : 1;
    res->its = rep.its; res->reason = rep.reason; res->inner_its = rep.inner_its; res->nlevels = rep.nlevels;
    res->fnorm0 = rep.fnorm0; res->fnorm1 = rep.fnorm1; res->rnorm0 = rep.rnorm0; res->rnorm = rep.rnorm;
    res->true_relres = rep.true_relres; res->setup_ms = rep.setup_ms; res->solve_ms = rep.solve_ms;
    res->emax_fine = rep.emax_fine;
    errors(&S, ua, &res->errinf, &res->err2h);
    const long n = (long)res->mx * res->my * res->mz;
    if (u_out) memcpy(u_out, ua, sizeof(double) * n);
    char gridstr[99];
    if (p->dim == 1) snprintf(gridstr, 99, "%d point 1D", p->mx);
    else if (p->dim == 2) snprintf(gridstr, 99, "%d x %d point 2D", p->mx, p->my);
    else snprintf(gridstr, 99, "%d x %d x %d point 3D", p->mx, p->my, p->mz);
    outf(&o, "problem %s on %s grid:\n  error |u-uexact|_inf = %.3e, |u-uexact|_h = %.3e\n",
         fish_problem_names[p->problem], gridstr, res->errinf, res->err2h);
    DMRestoreGlobalVector(S.da, &u);
    setup_free(&S);
    return 0;
}

/* ----------------------------------------------------------------- session */

struct fish_session_s {
    Setup   S;
    double *u0, *y;       /* host */
    double *d_b, *d_y;    /* device: F(u0) and the KSP solution */
    long    n;
};

int fish_write_operator(int argc, const char **argv, const char *path) {
    Setup S;
    char err[1200];
    int e = setup(argc, argv, &S, err, sizeof err, 0);
    if (e) {
        setup_free(&S);
        return p4g_seterr(e, "%s", err);
    }
    e = p4g_da_mat_view_binary(S.da, path, 0);
    setup_free(&S);
    return e;
}

int fish_session_create(int argc, const char **argv, fish_session *out) {
    struct fish_session_s *fs = (struct fish_session_s *)calloc(1, sizeof *fs);
    char err[1200];
    int e = setup(argc, argv, &fs->S, err, sizeof err, 1);
    if (e) {
        setup_free(&fs->S);
        free(fs);
        return p4g_seterr(e, "%s", err);
    }
    const FishProblem *p = &fs->S.ctx.pr;
    fs->n = (long)p->mx * (p->dim > 1 ? p->my : 1) * (p->dim > 2 ? p->mz : 1);
    fs->u0 = (double *)calloc(fs->n, sizeof(double));
    fs->y = (double *)calloc(fs->n, sizeof(double));
    double *F = (double *)calloc(fs->n, sizeof(double));
    Vec u;
    DMGetGlobalVector(fs->S.da, &u);
    e = InitialState(fs->S.da, fs->S.initial_random ? RANDOM : ZEROS,
                     fs->S.gonboundary ? PETSC_TRUE : PETSC_FALSE, u, &fs->S.ctx.pc);
    PetscScalar *ua;
    VecGetArray(u, &ua);
    memcpy(fs->u0, ua, sizeof(double) * fs->n);
    DMRestoreGlobalVector(fs->S.da, &u);
    if (!e) {
        /* F(u0) by the residual callback on pointer views of host arrays */
        DMDALocalInfo info;
        DMDAGetLocalInfo(fs->S.da, &info);
        Vec vu, vf;
        DMCreateGlobalVector(fs->S.da, &vu);
        DMCreateGlobalVector(fs->S.da, &vf);
        PetscScalar *pu, *pf;
        VecGetArray(vu, &pu);
        memcpy(pu, fs->u0, sizeof(double) * fs->n);
        void *au, *af;
        DMDAVecGetArray(fs->S.da, vu, &au);
        DMDAVecGetArray(fs->S.da, vf, &af);
        DMDASNESFunctionFn *fn = fs->S.ctx.pr.problem == PR_VARCOEF
                                     ? (DMDASNESFunctionFn *)&VarPoisson3DFunctionLocal
                                     : residual_ptr[p->dim - 1];
        e = fn(&info, au, af, &fs->S.ctx.pc);
        VecGetArray(vf, &pf);
        memcpy(F, pf, sizeof(double) * fs->n);
        VecDestroy(&vu);
        VecDestroy(&vf);
    }
    if (!e) e = p4g_solver_setup(fs->S.solver);
    if (!e) {
        fs->d_b = (double *)p4g_device_malloc(sizeof(double) * fs->n);
        fs->d_y = (double *)p4g_device_malloc(sizeof(double) * fs->n);
        if (!fs->d_b || !fs->d_y) e = PETSC_ERR_MEM;
    }
    if (!e) e = p4g_memcpy_htod(fs->d_b, F, sizeof(double) * fs->n);
    free(F);
    if (e) {
        fish_session_destroy(&fs);
        return e;
    }
    *out = fs;
    return 0;
}

int fish_session_solve(fish_session s, p4g_Report *rep) {
    if (!s) return p4g_seterr(PETSC_ERR_ARG_WRONG, "fish_session_solve: null session");
    return p4g_ksp_solve_device(s->S.solver, s->d_b, s->d_y, rep);
}

int fish_session_solution(fish_session s, double *u) {
    if (!s || !u) return p4g_seterr(PETSC_ERR_ARG_WRONG, "fish_session_solution: null argument");
    int e = p4g_memcpy_dtoh(s->y, s->d_y, sizeof(double) * s->n);
    if (e) return e;
    for (long q = 0; q < s->n; q++) u[q] = s->u0[q] - s->y[q];
    return 0;
}

int fish_session_errors(fish_session s, double *einf, double *e2h) {
    if (!s) return p4g_seterr(PETSC_ERR_ARG_WRONG, "fish_session_errors: null session");
    double *u = (double *)malloc(sizeof(double) * s->n);
    int e = fish_session_solution(s, u);
    if (!e) errors(&s->S, u, einf, e2h);
    free(u);
    return e;
}

p4g_Solver fish_session_solver(fish_session s) { return s ? s->S.solver : NULL; }
const double *fish_session_rhs_device(fish_session s) { return s ? s->d_b : NULL; }

int fish_session_sizes(fish_session s, int *dim, int *mx, int *my, int *mz) {
    if (!s) return p4g_seterr(PETSC_ERR_ARG_WRONG, "fish_session_sizes: null session");
    const FishProblem *p = &s->S.ctx.pr;
    if (dim) *dim = p->dim;
    if (mx) *mx = p->mx;
    if (my) *my = p->dim > 1 ? p->my : 1;
    if (mz) *mz = p->dim > 2 ? p->mz : 1;
    return 0;
}

int fish_session_destroy(fish_session *ps) {
    if (!ps || !*ps) return 0;
    struct fish_session_s *s = *ps;
    p4g_device_free(s->d_b);
    p4g_device_free(s->d_y);
    free(s->u0);
    free(s->y);
    setup_free(&s->S);
    free(s);
    *ps = NULL;
    return 0;
}
