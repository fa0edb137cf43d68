/*
 * poissonfunctions.c -- residual/Jacobian callbacks (see poissonfunctions.h).
 *
 * Discretization of c/ch6/poissonfunctions.c:4-258: vertex grid over the
 * DM's bounding box, x_i = xmin + i h with h = (xmax-xmin)/(m-1); rows scaled by the cell measure (1D: 1/h, 2D: the
 * hx/hy ratios, 3D: dvol/h^2), so the diagonal is constant, boundary rows
 * included:  F = scdiag (u - g) on the boundary; in the interior
 * F = scdiag u - sum_dir sc_dir (u_- + u_+) - measure f, where a neighbour
 * on the boundary contributes g instead of u.  The Jacobian keeps only
 * couplings between interior vertices.
 */
#include <math.h>
#include "poissonfunctions.h"

typedef struct {
    double h[3];
    double lo[3];     /* lower corner of the box        */
    double sc[3];     /* coupling per direction        */
    double scdiag;    /* diagonal of every row          */
    double measure;   /* multiplies f in the residual   */
} Scales;

static PetscErrorCode scales(DMDALocalInfo *info, PoissonCtx *user, Scales *s) {
    PetscReal lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    PetscCall(DMGetBoundingBox(info->da, lo, hi));
    const int m[3] = {info->mx, info->my, info->mz};
    const double c[3] = {user->cx, user->cy, user->cz};
    for (int d = 0; d < 3; d++) s->h[d] = d < info->dim ? (hi[d] - lo[d]) / (m[d] - 1) : 1.0;
    for (int d = 0; d < 3; d++) s->lo[d] = d < info->dim ? lo[d] : 0.0;
    if (info->dim == 1) {
        s->sc[0] = c[0] / s->h[0];
        s->sc[1] = s->sc[2] = 0.0;
        s->scdiag = 2.0 * s->sc[0];
        s->measure = s->h[0];
    } else if (info->dim == 2) {
        s->sc[0] = c[0] * s->h[1] / s->h[0];
        s->sc[1] = c[1] * s->h[0] / s->h[1];
        s->sc[2] = 0.0;
        s->scdiag = 2.0 * (s->sc[0] + s->sc[1]);
        s->measure = s->h[0] * s->h[1];
    } else {
        const double vol = s->h[0] * s->h[1] * s->h[2];
        for (int d = 0; d < 3; d++) s->sc[d] = c[d] * vol / (s->h[d] * s->h[d]);
        s->scdiag = 2.0 * (s->sc[0] + s->sc[1] + s->sc[2]);
        s->measure = vol;
    }
    return 0;
}

/* neighbour value: g there if it lies on the boundary, else the iterate */
#define NBR(onbdry, gx, gy, gz, val) ((onbdry) ? user->g_bdry((gx), (gy), (gz), user) : (val))

PetscErrorCode Poisson1DFunctionLocal(DMDALocalInfo *info, PetscReal *au, PetscReal *aF,
                                      PoissonCtx *user) {
    Scales s;
    PetscCall(scales(info, user, &s));
    const double h = s.h[0];
    for (int i = info->xs; i < info->xs + info->xm; i++) {
        const double x = s.lo[0] + i * h;
        if (i == 0 || i == info->mx - 1) {
            aF[i] = s.scdiag * (au[i] - user->g_bdry(x, 0.0, 0.0, user));
            continue;
        }
        const double uw = NBR(i - 1 == 0, x - h, 0.0, 0.0, au[i - 1]);
        const double ue = NBR(i + 1 == info->mx - 1, x + h, 0.0, 0.0, au[i + 1]);
        aF[i] = s.sc[0] * (2.0 * au[i] - uw - ue) - s.measure * user->f_rhs(x, 0.0, 0.0, user);
    }
    PetscCall(PetscLogFlops(9.0 * info->xm));
    return 0;
}

PetscErrorCode Poisson2DFunctionLocal(DMDALocalInfo *info, PetscReal **au, PetscReal **aF,
                                      PoissonCtx *user) {
    Scales s;
    PetscCall(scales(info, user, &s));
    const double hx = s.h[0], hy = s.h[1];
    for (int j = info->ys; j < info->ys + info->ym; j++) {
        const double y = s.lo[1] + j * hy;
        for (int i = info->xs; i < info->xs + info->xm; i++) {
            const double x = s.lo[0] + i * hx;
            if (i == 0 || i == info->mx - 1 || j == 0 || j == info->my - 1) {
                aF[j][i] = s.scdiag * (au[j][i] - user->g_bdry(x, y, 0.0, user));
                continue;
            }
            const double uw = NBR(i - 1 == 0, x - hx, y, 0.0, au[j][i - 1]);
            const double ue = NBR(i + 1 == info->mx - 1, x + hx, y, 0.0, au[j][i + 1]);
            const double us = NBR(j - 1 == 0, x, y - hy, 0.0, au[j - 1][i]);
            const double un = NBR(j + 1 == info->my - 1, x, y + hy, 0.0, au[j + 1][i]);
            aF[j][i] = s.scdiag * au[j][i] - s.sc[0] * (uw + ue) - s.sc[1] * (us + un)
                       - s.measure * user->f_rhs(x, y, 0.0, user);
        }
    }
    PetscCall(PetscLogFlops(11.0 * info->xm * info->ym));
    return 0;
}

PetscErrorCode Poisson3DFunctionLocal(DMDALocalInfo *info, PetscReal ***au, PetscReal ***aF,
                                      PoissonCtx *user) {
    Scales s;
    PetscCall(scales(info, user, &s));
    const double hx = s.h[0], hy = s.h[1], hz = s.h[2];
    for (int k = info->zs; k < info->zs + info->zm; k++) {
        const double z = s.lo[2] + k * hz;
        for (int j = info->ys; j < info->ys + info->ym; j++) {
            const double y = s.lo[1] + j * hy;
            for (int i = info->xs; i < info->xs + info->xm; i++) {
                const double x = s.lo[0] + i * hx;
                if (i == 0 || i == info->mx - 1 || j == 0 || j == info->my - 1 || k == 0 ||
                    k == info->mz - 1) {
                    aF[k][j][i] = s.scdiag * (au[k][j][i] - user->g_bdry(x, y, z, user));
                    continue;
                }
                const double uw = NBR(i - 1 == 0, x - hx, y, z, au[k][j][i - 1]);
                const double ue = NBR(i + 1 == info->mx - 1, x + hx, y, z, au[k][j][i + 1]);
                const double us = NBR(j - 1 == 0, x, y - hy, z, au[k][j - 1][i]);
                const double un = NBR(j + 1 == info->my - 1, x, y + hy, z, au[k][j + 1][i]);
                const double ud = NBR(k - 1 == 0, x, y, z - hz, au[k - 1][j][i]);
                const double uu = NBR(k + 1 == info->mz - 1, x, y, z + hz, au[k + 1][j][i]);
                aF[k][j][i] = s.scdiag * au[k][j][i] - s.sc[0] * (uw + ue) - s.sc[1] * (us + un)
                              - s.sc[2] * (ud + uu) - s.measure * user->f_rhs(x, y, z, user);
            }
        }
    }
    PetscCall(PetscLogFlops(14.0 * info->xm * info->ym * info->zm));
    return 0;
}

/* one Jacobian row: diagonal, plus couplings to interior neighbours of an
 * interior vertex (k,j,i) */
static PetscErrorCode jac_row(DMDALocalInfo *info, Mat Jpre, const Scales *s, int i, int j,
                              int k) {
    MatStencil row, col[7];
    double v[7];
    int n = 0;
    const int dim = info->dim;
    row.i = i; row.j = j; row.k = k; row.c = 0;
    col[0] = row;
    v[n++] = s->scdiag;
    int inside = i > 0 && i < info->mx - 1;
    if (dim > 1) inside = inside && j > 0 && j < info->my - 1;
    if (dim > 2) inside = inside && k > 0 && k < info->mz - 1;
    if (inside) {
        const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        const int m[3] = {info->mx, info->my, info->mz};
        for (int q = 0; q < 2 * dim; q++) {
            const int d = q / 2;
            const int p[3] = {i + off[q][0], j + off[q][1], k + off[q][2]};
            if (p[d] == 0 || p[d] == m[d] - 1) continue;   /* boundary neighbour */
            col[n] = row;
            col[n].i = p[0]; col[n].j = p[1]; col[n].k = p[2];
            v[n++] = -s->sc[d];
        }
    }
    return MatSetValuesStencil(Jpre, 1, &row, n, col, v, INSERT_VALUES);
}

static PetscErrorCode jacobian(DMDALocalInfo *info, Mat J, Mat Jpre, PoissonCtx *user) {
    Scales s;
    PetscCall(scales(info, user, &s));
    for (int k = info->zs; k < info->zs + (info->dim > 2 ? info->zm : 1); k++)
        for (int j = info->ys; j < info->ys + (info->dim > 1 ? info->ym : 1); j++)
            for (int i = info->xs; i < info->xs + info->xm; i++)
                PetscCall(jac_row(info, Jpre, &s, i, j, k));
    PetscCall(MatAssemblyBegin(Jpre, MAT_FINAL_ASSEMBLY));
    PetscCall(MatAssemblyEnd(Jpre, MAT_FINAL_ASSEMBLY));
    if (J != Jpre) {
        PetscCall(MatAssemblyBegin(J, MAT_FINAL_ASSEMBLY));
        PetscCall(MatAssemblyEnd(J, MAT_FINAL_ASSEMBLY));
    }
    return 0;
}

PetscErrorCode Poisson1DJacobianLocal(DMDALocalInfo *info, PetscReal *au, Mat J, Mat Jpre,
                                      PoissonCtx *user) {
    (void)au;
    return jacobian(info, J, Jpre, user);
}
PetscErrorCode Poisson2DJacobianLocal(DMDALocalInfo *info, PetscReal **au, Mat J, Mat Jpre,
                                      PoissonCtx *user) {
    (void)au;
    return jacobian(info, J, Jpre, user);
}
PetscErrorCode Poisson3DJacobianLocal(DMDALocalInfo *info, PetscReal ***au, Mat J, Mat Jpre,
                                      PoissonCtx *user) {
    (void)au;
    return jacobian(info, J, Jpre, user);
}

/* initial iterate: zeros or uniform random, then g on the boundary unless
 * gbdry is false (c/ch6/poissonfunctions.c:260-346) */
PetscErrorCode InitialState(DM da, InitialType it, PetscBool gbdry, Vec u, PoissonCtx *user) {
    DMDALocalInfo info;
    if (it == ZEROS) {
        PetscCall(VecSet(u, 0.0));
    } else if (it == RANDOM) {
        PetscRandom r;
        PetscCall(PetscRandomCreate(PETSC_COMM_WORLD, &r));
        PetscCall(VecSetRandom(u, r));
        PetscCall(PetscRandomDestroy(&r));
    } else {
        SETERRQ(PETSC_COMM_SELF, 4, "invalid InitialType ... how did I get here?\n");
    }
    if (!gbdry) return 0;
    PetscCall(DMDAGetLocalInfo(da, &info));
    PetscReal lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    PetscCall(DMGetBoundingBox(da, lo, hi));
    PetscScalar *a;
    PetscCall(VecGetArray(u, &a));
    const int mx = info.mx, my = info.dim > 1 ? info.my : 1, mz = info.dim > 2 ? info.mz : 1;
    double h[3] = {1, 1, 1};
    for (int d = 0; d < info.dim; d++) h[d] = (hi[d] - lo[d]) / ((d == 0 ? mx : d == 1 ? my : mz) - 1);
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int b = i == 0 || i == mx - 1;
                if (info.dim > 1) b = b || j == 0 || j == my - 1;
                if (info.dim > 2) b = b || k == 0 || k == mz - 1;
                if (b)
                    a[((long)k * my + j) * mx + i] =
                        user->g_bdry(lo[0] + i * h[0], info.dim > 1 ? lo[1] + j * h[1] : 0.0,
                                     info.dim > 2 ? lo[2] + k * h[2] : 0.0, user);
            }
    PetscCall(VecRestoreArray(u, &a));
    return 0;
}

/* -div(k grad u) - f: face coefficient = mean of the 4 cells sharing the
 * face's edge (times dvol/h^2); boundary rows: 2(ax+ay+az) times the mean
 * of the cells touching the vertex, times (u - g). */
PetscErrorCode VarPoisson3DFunctionLocal(DMDALocalInfo *info, PetscReal ***au, PetscReal ***aF,
                                         PoissonCtx *user) {
    const VarCoefCtx *vc = (const VarCoefCtx *)user->addctx;
    Scales s;
    PetscCall(scales(info, user, &s));
    const double hx = s.h[0], hy = s.h[1], hz = s.h[2], vol = hx * hy * hz;
    const double ax = vol / (hx * hx), ay = vol / (hy * hy), az = vol / (hz * hz);
    const int cx = info->mx - 1, cy = info->my - 1, cz = info->mz - 1;
#define KC(a, b, c) vc->kcell[((long)(c) * cy + (b)) * cx + (a)]
    for (int k = info->zs; k < info->zs + info->zm; k++)
        for (int j = info->ys; j < info->ys + info->ym; j++)
            for (int i = info->xs; i < info->xs + info->xm; i++) {
                const double x = s.lo[0] + i * hx, y = s.lo[1] + j * hy, z = s.lo[2] + k * hz;
                if (i == 0 || i == info->mx - 1 || j == 0 || j == info->my - 1 || k == 0 ||
                    k == info->mz - 1) {
                    double sum = 0.0;
                    int cnt = 0;
                    for (int dk = -1; dk <= 0; dk++)
                        for (int dj = -1; dj <= 0; dj++)
                            for (int di = -1; di <= 0; di++) {
                                const int a = i + di, b = j + dj, c = k + dk;
                                if (a < 0 || b < 0 || c < 0 || a >= cx || b >= cy || c >= cz) continue;
                                sum += KC(a, b, c);
                                cnt++;
                            }
                    aF[k][j][i] = (au[k][j][i] - user->g_bdry(x, y, z, user)) * (2.0 * (ax + ay + az) * (sum / cnt));
                    continue;
                }
                const double kw = 0.25 * (KC(i - 1, j - 1, k - 1) + KC(i - 1, j, k - 1) + KC(i - 1, j - 1, k) + KC(i - 1, j, k)) * ax;
                const double ke = 0.25 * (KC(i, j - 1, k - 1) + KC(i, j, k - 1) + KC(i, j - 1, k) + KC(i, j, k)) * ax;
                const double ks = 0.25 * (KC(i - 1, j - 1, k - 1) + KC(i, j - 1, k - 1) + KC(i - 1, j - 1, k) + KC(i, j - 1, k)) * ay;
                const double kn = 0.25 * (KC(i - 1, j, k - 1) + KC(i, j, k - 1) + KC(i - 1, j, k) + KC(i, j, k)) * ay;
                const double kd = 0.25 * (KC(i - 1, j - 1, k - 1) + KC(i, j - 1, k - 1) + KC(i - 1, j, k - 1) + KC(i, j, k - 1)) * az;
                const double ku = 0.25 * (KC(i - 1, j - 1, k) + KC(i, j - 1, k) + KC(i - 1, j, k) + KC(i, j, k)) * az;
                const double uw = NBR(i - 1 == 0, x - hx, y, z, au[k][j][i - 1]);
                const double ue = NBR(i + 1 == info->mx - 1, x + hx, y, z, au[k][j][i + 1]);
                const double us = NBR(j - 1 == 0, x, y - hy, z, au[k][j - 1][i]);
                const double un = NBR(j + 1 == info->my - 1, x, y + hy, z, au[k][j + 1][i]);
                const double ud = NBR(k - 1 == 0, x, y, z - hz, au[k - 1][j][i]);
                const double uu = NBR(k + 1 == info->mz - 1, x, y, z + hz, au[k + 1][j][i]);
                const double dg = kw + ke + ks + kn + kd + ku;
                aF[k][j][i] = dg * au[k][j][i] - kw * uw - ke * ue - ks * us - kn * un - kd * ud - ku * uu
                              - vol * user->f_rhs(x, y, z, user);
            }
#undef KC
    PetscCall(PetscLogFlops(40.0 * info->xm * info->ym * info->zm));
    return 0;
}
