/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).  Internal types.
 */
#ifndef P4G_ORACLE_INTERNAL_H
#define P4G_ORACLE_INTERNAL_H

#include <stdint.h>
#include <stdio.h>
#include "oracle.h"

/* ---- problems (c/ch6/fish.c:15-82,108-123, plus the sine, sine-mode and
 *      variable-coefficient problems of the benchmark configurations) ---- */
enum { PROB_MANUPOLY = 0, PROB_MANUEXP, PROB_ZERO, PROB_MANUSIN,
       PROB_SINEMODES, PROB_VARCOEF, PROB_COUNT };
extern const char *oracle_problem_names[PROB_COUNT];

#define NMODES 8
#define NFEAT 32

typedef struct {
    int    dim, problem;
    double Lx, Ly, Lz, cx, cy, cz;
    /* additive noise on f at interior nodes (Philox normal, node index) */
    double noise_sigma; uint64_t noise_seed;
    int    mx, my, mz;              /* finest grid, for node indices  */
    /* sine modes: f = sum a_m * pi^2 (cx p^2 + cy q^2 + cz r^2) mode_m */
    uint64_t modes_seed;
    int    mp[NMODES], mq[NMODES], mr[NMODES];
    double ma[NMODES];
    /* variable coefficient k = exp(g), g random Fourier features */
    uint64_t k_seed;
    double kw[NFEAT][3], kphi[NFEAT], ksig, kell;
} oprob;

void   oprob_init_tables(oprob *p);
double oprob_g(const oprob *p, double x, double y, double z);
double oprob_f(const oprob *p, double x, double y, double z);
double oprob_k(const oprob *p, double x, double y, double z);

/* ---- CSR matrix ---- */
typedef struct {
    long     n;
    int64_t *rp;
    int32_t *ci;
    double  *v;
} ocsr;
void ocsr_free(ocsr *A);
void ocsr_mult(const ocsr *A, const double *x, double *y);           /* y = A x */
void ocsr_resid(const ocsr *A, const double *b, const double *x, double *r);
int  ocsr_is_symmetric(const ocsr *A, double tol);

/* ---- grids ---- */
typedef struct {
    int  dim, mx, my, mz;
    long n;
    double Lx, Ly, Lz, hx, hy, hz;
} ogrid;
void ogrid_set(ogrid *g, int dim, int mx, int my, int mz, double Lx, double Ly,
               double Lz);
int  ogrid_coarsen(const ogrid *f, ogrid *c);        /* 0 ok, else error */

/* residual F(u) (poissonfunctions.c:4-115) and Jacobian (:117-258) */
void oracle_residual(const ogrid *g, const oprob *p, const double *u, double *F);
void oracle_jacobian(const ogrid *g, const oprob *p, ocsr *A);
/* variable coefficient: cell field, residual, Jacobian, coarse averaging */
void oracle_kcell_fill(const ogrid *g, const oprob *p, double *kcell);
void oracle_kcell_coarsen(const ogrid *f, const double *kf, double *kc);
void oracle_varcoef_jacobian(const ogrid *g, const double *kcell, ocsr *A);
void oracle_varcoef_residual(const ogrid *g, const oprob *p, const double *kcell,
                             const double *u, double *F);
void oracle_initial_state(const ogrid *g, const oprob *p, int random, int gbdry,
                          double *u);
void oracle_uexact(const ogrid *g, const oprob *p, double *u);

/* Q1 transfers */
void oq1_interp_add(const ogrid *c, const ogrid *f, const double *xc, double *xf);
void oq1_restrict(const ogrid *f, const ogrid *c, const double *rf, double *bc);
void oq1_build_P(const ogrid *c, const ogrid *f, ocsr *P);

/* ---- vectors (OpenMP) ---- */
double ovdot(long n, const double *x, const double *y);
double ovnorm2(long n, const double *x);
double ovnorminf(long n, const double *x);
void   ovcopy(long n, const double *x, double *y);
void   ovset(long n, double a, double *x);

/* ---- options ---- */
typedef struct { int n; const char *key[256]; const char *val[256]; int used[256]; } oopts;
void        oopts_parse(oopts *o, int argc, const char **argv);
const char *oopts_get(oopts *o, const char *key);          /* NULL if absent */
int         oopts_has(oopts *o, const char *key);
int         oopts_int(oopts *o, const char *key, int def);
double      oopts_real(oopts *o, const char *key, double def);
int         oopts_bool(oopts *o, const char *key, int def);

/* ---- solvers (oracle_solver.c) ---- */
enum { KSP_CG = 0, KSP_RICHARDSON, KSP_CHEBYSHEV, KSP_PREONLY, KSP_GMRES };
enum { PC_NONE = 0, PC_JACOBI, PC_SOR, PC_ILU, PC_LU, PC_MG, PC_SHELL };
enum { NORM_NONE = 0, NORM_PRECONDITIONED, NORM_UNPRECONDITIONED, NORM_NATURAL };

/* KSPConvergedReason values */
#define R_CONVERGED_RTOL 2
#define R_CONVERGED_ATOL 3
#define R_CONVERGED_ITS 4
#define R_DIVERGED_NULL (-2)
#define R_DIVERGED_ITS (-3)
#define R_DIVERGED_DTOL (-4)
#define R_DIVERGED_BREAKDOWN (-5)
#define R_DIVERGED_INDEFINITE_PC (-8)
#define R_DIVERGED_NANORINF (-9)
#define R_DIVERGED_INDEFINITE_MAT (-10)
const char *oracle_reason_name(int reason);

struct opc;
typedef struct omg omg;

typedef struct opc {
    int          type;
    const ocsr  *A;
    double      *idiag;               /* jacobi, sor                       */
    double       omega;               /* sor                               */
    int          nblk; long blk[17];  /* sor: row blocks (local sweeps)    */
    ocsr         F;  int64_t *dpos;   /* ilu: factors and diagonal offsets */
    double      *band; int bw;        /* lu: banded factor                 */
    omg         *mg;
    void       (*shell)(void *ctx, const double *r, double *z);   /* PCSHELL */
    void        *shell_ctx;
} opc;
void opc_setup(opc *pc, int type, const ocsr *A);
void opc_apply(opc *pc, const double *r, double *z);
void opc_free(opc *pc);
/* symmetric SOR sweeps with nonzero initial guess (PCApplyRichardson_SOR) */
void opc_sor_richardson(opc *pc, const double *b, double *x, int its);

typedef struct {
    int     type, pctype, maxit, normtype, monitor, guess_nonzero;
    int     restart;         /* gmres: Krylov dimension per cycle (30)  */
    double  rtol, atol, dtol, rich_scale;
    double  emin, emax;      /* chebyshev, after transform              */
    int     est_done;
    int     est_lanczos;     /* 1: Lanczos(CG), 0: GMRES Hessenberg     */
    int     esteig_steps;
    double  est_tf[4];       /* emin = tf0*min+tf1*max, emax = tf2*min+tf3*max */
    const long *gidx;        /* estimate start vector: noisy(gidx[i]), else noisy(i) */
    opc     pc;
    const ocsr *A;           /* the operator, or the PC's matrix when amult is set */
    void  (*amult)(void *ctx, const double *x, double *y);   /* shell operator (Amat != Pmat) */
    void   *actx;
    double *w[5];
    FILE   *mon; char *mbuf; int mlen; int *mpos; const char *prefix;
} oksp;
void oksp_init(oksp *k, int type, const ocsr *A, int pctype);
int  oracle_parse_ksp(const char *s, int def);   /* -1 when unknown */
int  oracle_parse_pc(const char *s, int def);
void oksp_free(oksp *k);
/* Solve A x = b; returns its; sets *reason, *rnorm0, *rnorm. */
int  oksp_solve(oksp *k, const double *b, double *x, int *reason, double *rnorm0,
                double *rnorm);
/* Chebyshev eigenvalue estimate (KSPChebyshev "esteig"). */
void oksp_estimate_eigs(oksp *k);

/* multigrid (PCMG) */
struct omg {
    int     nlev, cycles, galerkin;
    int     full;               /* -pc_mg_type full: one F-cycle per apply */
    int     kaskade;            /* -pc_mg_type kaskade                     */
    int     mcycles;            /* -pc_mg_multiplicative_cycles            */
    double  ttol, atol;         /* PCApplyRichardson_MG: stop test after the */
    int    *rreason;            /*   finest pre-smoothing (NULL: off)     */
    ogrid  *g;                  /* g[0] coarsest ... g[nlev-1] finest   */
    ocsr   *A;                  /* per level operators                  */
    oksp   *smooth;             /* per level smoother (0: coarse solve) */
    double **x, **b, **r;
    ocsr   *P;                  /* P[l]: level l-1 -> level l (galerkin) */
    int     nranks;             /* simulated MPI ranks (local SOR blocks) */
    int   (*split)[17];         /* per level ownership of the slowest dim */
};
int  omg_setup(omg *mg, const ogrid *fine, const oprob *prob, const ocsr *Afine,
               const double *kcell_fine, int nlev, oopts *o, char *err, int errlen);
void omg_apply(omg *mg, const double *b, double *x);
void omg_free(omg *mg);

/* A_c = P^T A P (CSR triple product, sorted columns) */
void oracle_galerkin(const ocsr *A, const ocsr *P, long nc, ocsr *Ac);

/* emit text like PetscPrintf into a growing buffer and/or stdout */
void oprintf(char *buf, int len, int *pos, const char *fmt, ...);

#endif
