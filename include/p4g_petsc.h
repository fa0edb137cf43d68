/*
 * p4g_petsc.h -- the PETSc-shaped host surface of the p4g library.
 *
 * DMDA-local residual and Jacobian callbacks written for PETSc, with the
 * signatures of c/ch6/poissonfunctions.h:56-84 (registered in the reference
 * by DMDASNESSetFunctionLocal / DMDASNESSetJacobianLocal, c/ch6/fish.c:225-228),
 * compile unchanged against this header and are called by p4g at set-up:
 * the residual once for F(u0), the Jacobian once per multigrid level, whose
 * rows are recorded and compressed into the device stencil.  Nothing here
 * runs inside the Krylov loop.
 *
 * Only the PETSc symbols the reference path's callbacks and driver use are
 * provided (c/ch6/poissonfunctions.c, c/ch6/fish.c).  All arrays are host
 * memory.  Index conventions are PETSc's: global (k,j,i) indices on
 * pointer-of-pointer views, x fastest.
 */
#ifndef P4G_PETSC_H
#define P4G_PETSC_H

#include <math.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int    PetscErrorCode;
typedef int    PetscInt;
typedef double PetscReal;
typedef double PetscScalar;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef int    MPI_Comm;
typedef int    PetscEnum;

#define PETSC_COMM_WORLD 0
#define PETSC_COMM_SELF  0
#define PETSC_DECIDE     (-1)

/* PETSc error numbers used by p4g */
#define PETSC_ERR_MEM       55
#define PETSC_ERR_SUP       56
#define PETSC_ERR_ARG_WRONG 62
#define PETSC_ERR_ARG_SIZ   60
#define PETSC_ERR_ARG_OUTOFRANGE 63
#define PETSC_ERR_ORDER     73
#define PETSC_ERR_LIB       76
#define PETSC_ERR_FILE_OPEN 65
#define PETSC_ERR_FILE_READ 66
#define PETSC_ERR_FILE_WRITE 67
#define PETSC_ERR_FILE_UNEXPECTED 79
#define PETSC_ERR_NOT_CONVERGED 91

typedef struct p4g_da_s     *DM;     /* a DMDA: structured vertex grid   */
typedef struct p4g_mat_s    *Mat;    /* recording matrix (set-up only)   */
typedef struct p4g_vec_s    *Vec;    /* host global vector of a DM       */
typedef struct p4g_random_s *PetscRandom;

typedef struct { PetscInt k, j, i, c; } MatStencil;
typedef enum { NOT_SET_VALUES = 0, INSERT_VALUES = 1, ADD_VALUES = 2 } InsertMode;
typedef enum { MAT_FLUSH_ASSEMBLY = 1, MAT_FINAL_ASSEMBLY = 0 } MatAssemblyType;
typedef enum { NORM_1 = 0, NORM_2 = 1, NORM_FROBENIUS = 2, NORM_INFINITY = 3 } NormType;

/* DMDAGetLocalInfo fields read by the callbacks (poissonfunctions.c:8-90) */
typedef struct {
    DM       da;
    PetscInt dim, dof, sw;
    PetscInt mx, my, mz;          /* global grid size        */
    PetscInt xs, ys, zs;          /* owned box start         */
    PetscInt xm, ym, zm;          /* owned box size          */
    PetscInt gxs, gys, gzs;       /* ghosted box start       */
    PetscInt gxm, gym, gzm;       /* ghosted box size        */
} DMDALocalInfo;

/* callback types (the reference casts to these, fish.c:92-100) */
typedef PetscErrorCode DMDASNESFunctionFn(DMDALocalInfo *info, void *x, void *f, void *ctx);
typedef PetscErrorCode DMDASNESJacobianFn(DMDALocalInfo *info, void *x, Mat J, Mat Jpre,
                                          void *ctx);

/* error propagation: PetscCall returns the code; SETERRQ records a message */
PetscErrorCode p4g_seterr(PetscErrorCode code, const char *fmt, ...);
#define PetscCall(call)                                                       \
    do {                                                                      \
        PetscErrorCode p4g_ierr_ = (call);                                    \
        if (p4g_ierr_) return p4g_ierr_;                                      \
    } while (0)
#define SETERRQ(comm, code, ...) return p4g_seterr((code), __VA_ARGS__)
#define PetscFunctionBeginUser
#define PetscFunctionReturn(x) return (x)

#define PetscExpReal(x)  exp(x)
#define PetscSqrtReal(x) sqrt(x)
#define PetscSinReal(x)  sin(x)
#define PetscCosReal(x)  cos(x)
#define PetscAbsReal(x)  fabs(x)
#define PetscPowReal(x, y) pow((x), (y))
#define PETSC_PI 3.14159265358979323846264338327950288

/* grid (DMDA) queries */
PetscErrorCode DMGetBoundingBox(DM da, PetscReal gmin[], PetscReal gmax[]);
/* the grid's box [min, max] per dimension (default [0, L] from p4g_da_create);
 * replaces DMDASetUniformCoordinates (c/ch12/obstacle.c:95) */
PetscErrorCode DMDASetUniformCoordinates(DM da, PetscReal xmin, PetscReal xmax, PetscReal ymin,
                                         PetscReal ymax, PetscReal zmin, PetscReal zmax);
PetscErrorCode DMDAGetLocalInfo(DM da, DMDALocalInfo *info);
PetscErrorCode DMSetApplicationContext(DM da, void *ctx);
PetscErrorCode DMGetApplicationContext(DM da, void *ctx);

/* recording matrix */
PetscErrorCode MatSetValuesStencil(Mat A, PetscInt m, const MatStencil idxm[], PetscInt n,
                                   const MatStencil idxn[], const PetscScalar v[],
                                   InsertMode mode);
PetscErrorCode MatAssemblyBegin(Mat A, MatAssemblyType t);
PetscErrorCode MatAssemblyEnd(Mat A, MatAssemblyType t);

/* host vectors of a DM (InitialState and the error report of fish.c) */
PetscErrorCode DMCreateGlobalVector(DM da, Vec *v);
PetscErrorCode DMGetGlobalVector(DM da, Vec *v);
PetscErrorCode DMRestoreGlobalVector(DM da, Vec *v);
PetscErrorCode VecDestroy(Vec *v);
PetscErrorCode VecSet(Vec v, PetscScalar a);
PetscErrorCode VecAXPY(Vec y, PetscScalar a, Vec x);
PetscErrorCode VecNorm(Vec v, NormType t, PetscReal *nrm);
PetscErrorCode VecGetSize(Vec v, PetscInt *n);
PetscErrorCode VecGetArray(Vec v, PetscScalar **a);
PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a);
PetscErrorCode VecSetRandom(Vec v, PetscRandom r);
PetscErrorCode PetscRandomCreate(MPI_Comm comm, PetscRandom *r);
PetscErrorCode PetscRandomDestroy(PetscRandom *r);
/* pointer-of-pointer views with global indices: PetscReal *, **, *** */
PetscErrorCode DMDAVecGetArray(DM da, Vec v, void *array);
PetscErrorCode DMDAVecRestoreArray(DM da, Vec v, void *array);

/* logging (PetscLogFlops counts are accumulated per DM) and printing */
PetscErrorCode PetscLogFlops(PetscReal n);
PetscErrorCode PetscPrintf(MPI_Comm comm, const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
