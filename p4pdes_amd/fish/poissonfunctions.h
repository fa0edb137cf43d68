/*
 * poissonfunctions.h -- DMDA-local residual and Jacobian callbacks of the
 * Poisson problem  -cx u_xx - cy u_yy - cz u_zz = f,  u = g on the boundary,
 * with the interface of the reference (c/ch6/poissonfunctions.h:43-95): the
 * same PoissonCtx layout, names, signatures and row scaling (rows scaled by
 * the cell volume, Dirichlet rows kept in the system with the interior
 * diagonal, boundary neighbours eliminated), compiled against the p4g
 * PETSc-shaped surface.  A variable-coefficient residual is added for
 * -div(k grad u) = f with cell-centred k.
 */
#ifndef P4G_POISSONFUNCTIONS_H
#define P4G_POISSONFUNCTIONS_H

#include "petsc.h"

typedef struct {
    PetscReal Lx, Ly, Lz;   /* domain [0,Lx] x [0,Ly] x [0,Lz] */
    PetscReal cx, cy, cz;   /* coefficients of -c u_dd         */
    PetscReal (*f_rhs)(PetscReal x, PetscReal y, PetscReal z, void *ctx);
    PetscReal (*g_bdry)(PetscReal x, PetscReal y, PetscReal z, void *ctx);
    void *addctx;           /* problem-specific data             */
} PoissonCtx;

typedef enum { ZEROS, RANDOM } InitialType;

PetscErrorCode Poisson1DFunctionLocal(DMDALocalInfo *info, PetscReal *au, PetscReal *aF,
                                      PoissonCtx *user);
PetscErrorCode Poisson2DFunctionLocal(DMDALocalInfo *info, PetscReal **au, PetscReal **aF,
                                      PoissonCtx *user);
PetscErrorCode Poisson3DFunctionLocal(DMDALocalInfo *info, PetscReal ***au, PetscReal ***aF,
                                      PoissonCtx *user);
PetscErrorCode Poisson1DJacobianLocal(DMDALocalInfo *info, PetscReal *au, Mat J, Mat Jpre,
                                      PoissonCtx *user);
PetscErrorCode Poisson2DJacobianLocal(DMDALocalInfo *info, PetscReal **au, Mat J, Mat Jpre,
                                      PoissonCtx *user);
PetscErrorCode Poisson3DJacobianLocal(DMDALocalInfo *info, PetscReal ***au, Mat J, Mat Jpre,
                                      PoissonCtx *user);
PetscErrorCode InitialState(DM da, InitialType it, PetscBool gbdry, Vec u, PoissonCtx *user);

/* -div(k grad u) - f with k on cells; addctx points to a VarCoefCtx */
typedef struct {
    const double *kcell;    /* (mx-1)(my-1)(mz-1), x fastest */
    int mx, my, mz;
} VarCoefCtx;
PetscErrorCode VarPoisson3DFunctionLocal(DMDALocalInfo *info, PetscReal ***au, PetscReal ***aF,
                                         PoissonCtx *user);

#endif
