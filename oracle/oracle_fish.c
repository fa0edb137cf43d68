/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * The `fish` driver (c/ch6/fish.c:128-286) on the oracle's solvers:
 * options -> problem -> DMDA-like grid (3-point base, -da_refine) ->
 * InitialState -> SNES KSPONLY (F(u0), J, KSP solve J y = F(u0),
 * u = u0 - y) -> error report in fish's format.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle_internal.h"


static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

typedef struct {
    oopts  o;
    oprob  prob;
    ogrid  g;
    int    refine, initial_random, gonboundary;
    ocsr   A;
    double *kcell;
    oksp   ksp;
    double *u0, *F0, *y;
    char   err[256];
} osession;

static void session_free(osession *s) {
    oksp_free(&s->ksp);
    ocsr_free(&s->A);
    free(s->kcell); free(s->u0); free(s->F0); free(s->y);
}

/* returns 0 or a fish.c-style error code (message in s->err) */
static int session_setup(osession *s, int argc, const char **argv, char *out, int outlen, int *pos) {
    memset(s, 0, sizeof *s);
    oopts *o = &s->o;
    oopts_parse(o, argc, argv);
    oprob *p = &s->prob;
    /* fish defaults (fish.c:140-153) */
    p->dim = oopts_int(o, "fsh_dim", 2);
    p->Lx = oopts_real(o, "fsh_Lx", 1.0); p->Ly = oopts_real(o, "fsh_Ly", 1.0);
    p->Lz = oopts_real(o, "fsh_Lz", 1.0);
    p->cx = oopts_real(o, "fsh_cx", 1.0); p->cy = oopts_real(o, "fsh_cy", 1.0);
    p->cz = oopts_real(o, "fsh_cz", 1.0);
    p->problem = PROB_MANUEXP;
    const char *pr = oopts_get(o, "fsh_problem");
    if (pr) {
        p->problem = -1;
        for (int q = 0; q < PROB_COUNT; q++)
            if (!strcmp(pr, oracle_problem_names[q])) p->problem = q;
        if (p->problem < 0) { snprintf(s->err, sizeof s->err, "unknown -fsh_problem %s", pr); return 62; }
    }
    const char *it = oopts_get(o, "fsh_initial_type");
    s->initial_random = it && !strcmp(it, "random");
    if (it && strcmp(it, "random") && strcmp(it, "zeros")) {
        snprintf(s->err, sizeof s->err, "unknown -fsh_initial_type %s", it); return 62;
    }
    s->gonboundary = oopts_bool(o, "fsh_initial_gonboundary", 1);
    p->noise_sigma = oopts_real(o, "fsh_noise_sigma", 0.0);
    p->noise_seed = (uint64_t)oopts_int(o, "fsh_noise_seed", 0);
    p->modes_seed = (uint64_t)oopts_int(o, "fsh_modes_seed", 0);
    p->k_seed = (uint64_t)oopts_int(o, "fsh_k_seed", 0);
    /* argument checks (fish.c:188-193, 214-215) */
    if (p->cx <= 0.0 || p->cy <= 0.0 || p->cz <= 0.0) {
        snprintf(s->err, sizeof s->err, "positivity required for coefficients cx,cy,cz\n"); return 2;
    }
    if (p->problem == PROB_MANUEXP && (p->cx != 1.0 || p->cy != 1.0 || p->cz != 1.0)) {
        snprintf(s->err, sizeof s->err, "cx=cy=cz=1 required for problem MANUEXP\n"); return 3;
    }
    if (p->dim < 1 || p->dim > 3) {
        snprintf(s->err, sizeof s->err, "invalid dim for DMDA creation\n"); return 1;
    }
    if (p->problem == PROB_VARCOEF && p->dim != 3) {
        snprintf(s->err, sizeof s->err, "problem varcoef requires -fsh_dim 3\n"); return 62;
    }
    oprob_init_tables(p);
    /* DMDA: 3-point base per dimension (fish.c:197-212), -da_grid_*, -da_refine */
    int m[3] = {oopts_int(o, "da_grid_x", 3), oopts_int(o, "da_grid_y", 3), oopts_int(o, "da_grid_z", 3)};
    s->refine = oopts_int(o, "da_refine", 0);
    for (int d = 0; d < 3; d++)
        for (int r = 0; r < s->refine; r++) m[d] = 2 * (m[d] - 1) + 1;
    ogrid_set(&s->g, p->dim, m[0], m[1], m[2], p->Lx, p->Ly, p->Lz);
    p->mx = s->g.mx; p->my = s->g.my; p->mz = s->g.mz;
    long n = s->g.n;
    s->u0 = (double *)malloc(sizeof(double) * n);
    s->F0 = (double *)malloc(sizeof(double) * n);
    s->y = (double *)calloc(n, sizeof(double));
    oracle_initial_state(&s->g, p, s->initial_random, s->gonboundary, s->u0);
    if (p->problem == PROB_VARCOEF) {
        s->kcell = (double *)malloc(sizeof(double) * (size_t)(s->g.mx - 1) * (s->g.my - 1) * (s->g.mz - 1));
        oracle_kcell_fill(&s->g, p, s->kcell);
        oracle_varcoef_residual(&s->g, p, s->kcell, s->u0, s->F0);
        oracle_varcoef_jacobian(&s->g, s->kcell, &s->A);
    } else {
        oracle_residual(&s->g, p, s->u0, s->F0);
        oracle_jacobian(&s->g, p, &s->A);
    }
    const char *sym = oopts_get(o, "mat_is_symmetric");
    if (sym) {
        double tol = *sym ? atof(sym) : 0.0;
        /* after each assembly: the preallocated zero matrix, then the Jacobian */
        oprintf(out, outlen, pos, "Matrix is symmetric (tolerance %g)\n", tol);
        oprintf(out, outlen, pos, "Matrix is %ssymmetric (tolerance %g)\n",
                ocsr_is_symmetric(&s->A, tol) ? "" : "not ", tol);
    }
    oopts_get(o, "snes_fd_color");   /* exact Jacobian is used (linear problem) */
    /* KSP (fish.c:231-234) */
    int ktype = oracle_parse_ksp(oopts_get(o, "ksp_type"), KSP_CG);
    int pctype = oracle_parse_pc(oopts_get(o, "pc_type"), PC_ILU);
    if (ktype < 0 || pctype < 0) { snprintf(s->err, sizeof s->err, "unsupported -ksp_type/-pc_type"); return 56; }
    oksp_init(&s->ksp, ktype, &s->A, pctype);
    oksp *k = &s->ksp;
    k->rtol = oopts_real(o, "ksp_rtol", 1e-5);
    k->atol = oopts_real(o, "ksp_atol", 1e-50);
    k->dtol = oopts_real(o, "ksp_divtol", 1e5);
    k->maxit = oopts_int(o, "ksp_max_it", 10000);
    k->rich_scale = oopts_real(o, "ksp_richardson_scale", 1.0);
    k->restart = oopts_int(o, "ksp_gmres_restart", 30);
    const char *nt = oopts_get(o, "ksp_norm_type");
    /* KSPRICHARDSON keeps CG's preconditioned default (the reference's study
     * scripts always pass -ksp_norm_type unpreconditioned, so the default is
     * unpinned); the outer Chebyshev loop here computes no norms */
    if (ktype == KSP_CHEBYSHEV) k->normtype = NORM_UNPRECONDITIONED;
    if (ktype == KSP_RICHARDSON) k->normtype = NORM_PRECONDITIONED;
    if (nt) {
        if (!strcmp(nt, "preconditioned")) k->normtype = NORM_PRECONDITIONED;
        else if (!strcmp(nt, "unpreconditioned")) k->normtype = NORM_UNPRECONDITIONED;
        else if (!strcmp(nt, "natural")) k->normtype = NORM_NATURAL;
        else if (!strcmp(nt, "none")) k->normtype = NORM_NONE;
    }
    if (pctype == PC_MG) {
        int nlev = oopts_int(o, "pc_mg_levels", s->refine + 1);
        k->pc.mg = (omg *)calloc(1, sizeof(omg));
        int e = omg_setup(k->pc.mg, &s->g, p, &s->A, s->kcell, nlev, o, s->err, sizeof s->err);
        if (e) return e;
    }
    return 0;
}

/* one SNES KSPONLY linear solve from y = 0; returns its */
static int session_solve(osession *s, fish_result *res) {
    oksp *k = &s->ksp;
    ovset(s->g.n, 0.0, s->y);
    int reason; double r0, rn;
    int its = oksp_solve(k, s->F0, s->y, &reason, &r0, &rn);
    res->its = its; res->reason = reason; res->rnorm0 = r0; res->rnorm = rn;
    return its;
}

int oracle_fish_run(int argc, const char **argv, fish_result *res, double *u_out, char *out,
                    int outlen) {
    osession s;
    int pos = 0;
    fish_result dummy;
    if (!res) res = &dummy;
    memset(res, 0, sizeof *res);
    if (out && outlen > 0) out[0] = 0;
    double t0 = now_ms();
    int e = session_setup(&s, argc, argv, out, outlen, &pos);
    if (e) {
        oprintf(out, outlen, &pos, "[0]PETSC ERROR: %s", s.err);
        session_free(&s);
        return e;
    }
    oopts *o = &s.o;
    ogrid *g = &s.g;
    long n = g->n;
    res->dim = g->dim; res->mx = g->mx; res->my = g->my; res->mz = g->mz;
    res->nlevels = s.ksp.pc.mg ? s.ksp.pc.mg->nlev : 0;
    if (s.ksp.pc.mg && s.ksp.pc.mg->nlev > 1) res->emax_fine = s.ksp.pc.mg->smooth[s.ksp.pc.mg->nlev - 1].emax;
    int monitor = oopts_has(o, "snes_monitor_short") || oopts_has(o, "snes_monitor");
    s.ksp.monitor = oopts_has(o, "ksp_monitor");
    s.ksp.mbuf = out; s.ksp.mlen = outlen; s.ksp.mpos = &pos; s.ksp.prefix = "    ";
    res->fnorm0 = ovnorm2(n, s.F0);
    if (monitor) {
        if (res->fnorm0 > 1e-9) oprintf(out, outlen, &pos, "%3d SNES Function norm %g\n", 0, res->fnorm0);
        else if (res->fnorm0 > 1e-11) oprintf(out, outlen, &pos, "%3d SNES Function norm %5.3e\n", 0, res->fnorm0);
        else oprintf(out, outlen, &pos, "%3d SNES Function norm < 1.e-11\n", 0);
    }
    double t1 = now_ms();
    res->setup_ms = t1 - t0;
    session_solve(&s, res);
    res->solve_ms = now_ms() - t1;
    if (s.ksp.pc.mg && s.ksp.pc.mg->nlev > 1) res->emax_fine = s.ksp.pc.mg->smooth[s.ksp.pc.mg->nlev - 1].emax;
    /* true residual of the linear solve */
    double *w = (double *)malloc(sizeof(double) * n);
    ocsr_resid(&s.A, s.F0, s.y, w);
    res->true_relres = res->fnorm0 > 0 ? ovnorm2(n, w) / res->fnorm0 : 0.0;
    if (oopts_has(o, "ksp_converged_reason")) {
        if (res->reason > 0)
            oprintf(out, outlen, &pos, "    Linear solve converged due to %s iterations %d\n",
                    oracle_reason_name(res->reason), res->its);
        else
            oprintf(out, outlen, &pos, "    Linear solve did not converge due to %s iterations %d\n",
                    oracle_reason_name(res->reason), res->its);
    }
    /* u = u0 - y (SNESSolve_KSPONLY) */
    double *u = (double *)malloc(sizeof(double) * n);
    for (long q = 0; q < n; q++) u[q] = s.u0[q] - s.y[q];
    if (s.kcell) oracle_varcoef_residual(g, &s.prob, s.kcell, u, w);
    else oracle_residual(g, &s.prob, u, w);
    res->fnorm1 = ovnorm2(n, w);
    if (monitor) {
        if (res->fnorm1 > 1e-9) oprintf(out, outlen, &pos, "%3d SNES Function norm %g\n", 1, res->fnorm1);
        else if (res->fnorm1 > 1e-11) oprintf(out, outlen, &pos, "%3d SNES Function norm %5.3e\n", 1, res->fnorm1);
        else oprintf(out, outlen, &pos, "%3d SNES Function norm < 1.e-11\n", 1);
    }
    if (u_out) memcpy(u_out, u, sizeof(double) * n);
    /* error report (fish.c:248-280) */
    oracle_uexact(g, &s.prob, w);
    for (long q = 0; q < n; q++) w[q] = u[q] - w[q];
    res->errinf = ovnorminf(n, w);
    double err2 = ovnorm2(n, w), nc;
    char gridstr[99];
    if (g->dim == 1) { nc = sqrt((double)(g->mx - 1)); snprintf(gridstr, 99, "%d point 1D", g->mx); }
    else if (g->dim == 2) {
        nc = sqrt((double)(g->mx - 1) * (g->my - 1));
        snprintf(gridstr, 99, "%d x %d point 2D", g->mx, g->my);
    } else {
        nc = sqrt((double)(g->mx - 1) * (g->my - 1) * (g->mz - 1));
        snprintf(gridstr, 99, "%d x %d x %d point 3D", g->mx, g->my, g->mz);
    }
    res->err2h = err2 / nc;
    oprintf(out, outlen, &pos, "problem %s on %s grid:\n  error |u-uexact|_inf = %.3e, |u-uexact|_h = %.3e\n",
            oracle_problem_names[s.prob.problem], gridstr, res->errinf, res->err2h);
    free(u); free(w);
    session_free(&s);
    return 0;
}

/* one application of the preconditioner fish's KSP uses (PCApply) */
int oracle_pc_apply(int argc, const char **argv, const double *r, double *z) {
    osession s;
    int pos = 0;
    int e = session_setup(&s, argc, argv, NULL, 0, &pos);
    if (!e) opc_apply(&s.ksp.pc, r, z);
    session_free(&s);
    return e;
}

/* y = A x with the operator fish assembles (MatMult) */
int oracle_mat_mult(int argc, const char **argv, const double *x, double *y) {
    osession s;
    int pos = 0;
    int e = session_setup(&s, argc, argv, NULL, 0, &pos);
    if (!e) ocsr_mult(&s.A, x, y);
    session_free(&s);
    return e;
}

/* F(u0) of fish's initial iterate: the right-hand side of the KSP solve */
int oracle_rhs(int argc, const char **argv, double *F) {
    osession s;
    int pos = 0;
    int e = session_setup(&s, argc, argv, NULL, 0, &pos);
    if (!e) ovcopy(s.g.n, s.F0, F);
    session_free(&s);
    return e;
}

int oracle_fish_time(int argc, const char **argv, int nthreads, int warmup, int nsolves,
                     double *times_ms, fish_result *res) {
    return oracle_fish_time_sample(argc, argv, nthreads, warmup, nsolves, times_ms, res,
                                   NULL, 0, NULL, NULL);
}

int oracle_fish_time_sample(int argc, const char **argv, int nthreads, int warmup,
                            int nsolves, double *times_ms, fish_result *res,
                            double *setup_ms, int sample_its, double *sample_ms,
                            int *sample_done) {
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    osession s;
    int pos = 0;
    double ts = now_ms();
    int e = session_setup(&s, argc, argv, NULL, 0, &pos);
    if (setup_ms) *setup_ms = now_ms() - ts;
    if (e) { session_free(&s); return e; }
    fish_result r;
    memset(&r, 0, sizeof r);
    for (int w = 0; w < warmup; w++) session_solve(&s, &r);
    for (int q = 0; q < nsolves; q++) {
        double t0 = now_ms();
        session_solve(&s, &r);
        times_ms[q] = now_ms() - t0;
    }
    r.dim = s.g.dim; r.mx = s.g.mx; r.my = s.g.my; r.mz = s.g.mz;
    if (res) *res = r;
    if (sample_its > 0) {
#ifdef _OPENMP
        omp_set_num_threads(1);
#endif
        int maxit = s.ksp.maxit;
        s.ksp.maxit = sample_its;
        fish_result r1;
        memset(&r1, 0, sizeof r1);
        double t0 = now_ms();
        int its = session_solve(&s, &r1);
        if (sample_ms) *sample_ms = now_ms() - t0;
        if (sample_done) *sample_done = its;
        s.ksp.maxit = maxit;
    }
    session_free(&s);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}
