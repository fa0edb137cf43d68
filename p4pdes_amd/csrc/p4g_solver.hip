// p4g_solver.hip -- host orchestration of the gfx950 CG + geometric
// multigrid Poisson solver behind the C ABI of include/p4g.h.
//
// Semantics follow what the reference path gets from PETSc with
// `-ksp_type cg -pc_type mg` (c/ch6/fish.c:231-239):
//   * KSPCG: preconditioned PCG; default test |B r_k| <= max(rtol |B r_0|, atol);
//     beta == 0 -> CONVERGED_ATOL; a sign change of beta = (z, r) ->
//     DIVERGED_INDEFINITE_PC; (p, A p) == 0 or a sign change of it ->
//     DIVERGED_INDEFINITE_MAT (before x += alpha p); |r| >= dtol |r_0| ->
//     DIVERGED_DTOL.
//   * PCMG multiplicative cycle: pre-smooth, r = b - A x, b_c = P^T r,
//     x_c = 0, recurse (twice for W), x += P x_c, post-smooth; one smoother
//     per level with a nonzero initial guess; coarsest level solved directly.
//   * KSPCHEBYSHEV smoother with Jacobi: first step x += s D^-1 (b - A x),
//     s = 2/(emax+emin), then the three-term recurrence; bounds
//     [0.1, 1.1] * lambda_max estimated by 10 PCG steps from a fixed
//     pseudo-random right-hand side.  KSPRICHARDSON: x += w D^-1 (b - A x).
// Coarse operators are re-discretized by calling the user's Jacobian
// callback on each coarsened grid (fish.c:7).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "p4g_internal.hpp"
#include "p4g_kernels.hpp"

using namespace p4g;

#define HIP_TRY(call)                                                                      \
    do {                                                                                   \
        hipError_t e__ = (call);                                                           \
        if (e__ != hipSuccess)                                                             \
            return set_error(PETSC_ERR_LIB, "%s: %s (%s:%d)", #call, hipGetErrorString(e__), \
                             __FILE__, __LINE__);                                          \
    } while (0)

#define LAUNCH_CHECK()                                                                     \
    do {                                                                                   \
        hipError_t e__ = hipGetLastError();                                                \
        if (e__ != hipSuccess)                                                             \
            return set_error(PETSC_ERR_LIB, "kernel launch: %s (%s:%d)", hipGetErrorString(e__), \
                             __FILE__, __LINE__);                                          \
    } while (0)

namespace p4g {

enum { PCT_MG = 0, PCT_JACOBI, PCT_NONE, PCT_ILU, PCT_ASM };
enum { SM_CHEB = 0, SM_RICH };
enum { NT_PRECOND = 0, NT_UNPRECOND, NT_NATURAL };

// planes per marching chunk: short chunks on small grids (latency-bound),
// longer ones on large grids (fewer re-read halo planes)
// marching chunk: one plane per thread on small grids (latency-bound, want
// parallelism), longer chunks on large ones (fewer halo re-reads)
static inline int kchf(int dim, int my, int mz) {
    const int slow = dim == 3 ? mz : (dim == 2 ? my : 1);
    return slow >= 400 ? 8 : (slow >= 100 ? 4 : 1);
}
constexpr int VBLOCKS = 2048;      // blocks of the grid-stride vector kernels
constexpr int MAX_COARSE = 2048;   // largest coarsest grid for the direct solve

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// precision-independent description of one multigrid level
// the progress flags of a level's pipelined 2D SOR sweeps (k_sor9_strips):
// one word per 64-row strip, a stall word, and the host's sweep counter
struct SorPipe {
    double* pk = nullptr;                 // the operator as k_sor9_pack records (refreshed with it)
    unsigned long long* flags = nullptr;
    int* err = nullptr;
    unsigned long long epoch = 0;
    long spin_limit = 1L << 24;           // polls before a wait gives up
    unsigned long long need_extra = 0;    // -p4g_sor_test_stall: a target no strip reaches
};

struct LevelInfo {
    int  dim = 3, mx = 0, my = 1, mz = 1;
    long n = 0;
    double h[3] = {0, 0, 0};
    bool var = false;
    double d = 0, db = 0, sx = 0, sy = 0, sz = 0;    // constant stencil
    std::vector<double> kcell;                       // variable coefficient cells
    int cx = 0, cy = 0, cz = 0;
    double ax = 0, ay = 0, az = 0;
    double emin = 0, emax = 0;                       // Chebyshev bounds
    bool gal = false;                                // Galerkin operator P^T A P in 27-point form
    int sor_nblk = 1;                                // SOR local sweeps: blocks of the slowest dimension
    int sor_blk[17] = {0};                           //   and their first indices
    SorPipe* pipe = nullptr;                         // 2D 27-plane SOR by pipelined strips, when set
};

// one symmetric SOR sweep z = SSOR(r) on a level (omega 1, zero initial
// guess): the hyperplanes forward then backward (k_sor_plane)
template <typename T>
static int sor_run(const LevelInfo& L, const double* g27, const T* r, T* z, hipStream_t st) {
    if constexpr (std::is_same<T, double>::value) {
        if (L.dim == 3 && !L.gal && !L.var && L.pipe && L.pipe->flags && !L.pipe->pk && L.sor_nblk <= 1) {
            // 3D 7-point level: pipelined strips, one launch per sweep direction
            SorPipe* P = L.pipe;
            const unsigned nwg = (unsigned)(((L.my + 63) / 64) * L.mz);
            const unsigned long long b0 = (++P->epoch) << 22;
            k_sor7_strips<double><<<nwg, 64, 0, st>>>(L.mx, L.my, L.mz, L.d, L.db, L.sx, L.sy, L.sz, r, z, P->flags,
                                                       b0, 0, P->err, P->spin_limit, P->need_extra);
            const unsigned long long b1 = (++P->epoch) << 22;
            k_sor7_strips<double><<<nwg, 64, 0, st>>>(L.mx, L.my, L.mz, L.d, L.db, L.sx, L.sy, L.sz, r, z, P->flags,
                                                       b1, 1, P->err, P->spin_limit, P->need_extra);
            LAUNCH_CHECK();
            return 0;
        }
        if (L.gal && L.dim == 2 && L.pipe && L.pipe->pk && L.sor_nblk <= 1 && g27) {   // one launch per direction
            const int ns = (L.my + 63) / 64;
            SorPipe* P = L.pipe;
            const unsigned long long b0 = (++P->epoch) << 22;
            k_sor9_strips<<<ns, 64, 0, st>>>(L.mx, L.my, P->pk, r, z, P->flags, b0, 0, P->err, P->spin_limit,
                                             P->need_extra);
            const unsigned long long b1 = (++P->epoch) << 22;
            k_sor9_strips<<<ns, 64, 0, st>>>(L.mx, L.my, P->pk, r, z, P->flags, b1, 1, P->err, P->spin_limit,
                                             P->need_extra);
            LAUNCH_CHECK();
            return 0;
        }
    }
    SorOp o;
    o.dim = L.dim; o.mx = L.mx; o.my = L.my; o.mz = L.mz; o.n = L.n;
    o.d = L.d; o.db = L.db; o.sx = L.sx; o.sy = L.sy; o.sz = L.sz;
    o.g = L.gal ? g27 : nullptr;
    if (L.var) return set_error(PETSC_ERR_SUP, "SOR smoothing of a variable-coefficient level is not implemented");
    if (L.gal && !g27) return set_error(PETSC_ERR_LIB, "internal: SOR on a Galerkin level without coefficients");
    o.nblk = L.sor_nblk;
    for (int q = 0; q < 17; ++q) o.blk[q] = L.sor_blk[q];
    const int wj = o.g ? 2 : 1, wk = o.g ? 4 : 1;
    const int np = (L.mx - 1) + wj * (L.dim >= 2 ? L.my - 1 : 0) + wk * (L.dim == 3 ? L.mz - 1 : 0) + 1;
    const long nt = L.dim == 3 ? (long)L.my * L.mz : (L.dim == 2 ? (long)L.my : 1L);
    const int gb = (int)((nt + 255) / 256);
    for (int p = 0; p < np; ++p) k_sor_plane<T><<<gb, 256, 0, st>>>(o, p, 0, r, z);
    for (int p = np - 1; p >= 0; --p) k_sor_plane<T><<<gb, 256, 0, st>>>(o, p, 1, r, z);
    LAUNCH_CHECK();
    return 0;
}

// ILU(0) pivots of a constant-operator level in lexicographic order, the
// IKJ arithmetic of a row: u_q = a_qq - l_qm a_mq over the lower interior
// neighbours m in column order, l_qm = a_qm / u_m
static void ilu_pivots(const LevelInfo& L, std::vector<double>& u) {
    u.assign(L.n, 0.0);
    const long sy = L.mx, sz = (long)L.mx * L.my;
    auto bnd = [&](int i, int j, int k) {
        return i == 0 || i == L.mx - 1 || (L.dim >= 2 && (j == 0 || j == L.my - 1)) ||
               (L.dim == 3 && (k == 0 || k == L.mz - 1));
    };
    for (int k = 0; k < L.mz; ++k)
        for (int j = 0; j < L.my; ++j)
            for (int i = 0; i < L.mx; ++i) {
                const long q = ((long)k * L.my + j) * L.mx + i;
                if (bnd(i, j, k)) { u[q] = L.db; continue; }
                double v = L.d;
                if (L.dim == 3 && !bnd(i, j, k - 1)) v -= ((-L.sz) / u[q - sz]) * (-L.sz);
                if (L.dim >= 2 && !bnd(i, j - 1, k)) v -= ((-L.sy) / u[q - sy]) * (-L.sy);
                if (!bnd(i - 1, j, k)) v -= ((-L.sx) / u[q - 1]) * (-L.sx);
                u[q] = v;
            }
}

// ILU(0) of a 27-point-form level with 7-point nonzeros (reduced VI steps):
// pivots into u, then z = (L U)^-1 r
static SorOp sor_op27(const LevelInfo& L, const double* g27) {
    SorOp o;
    o.dim = L.dim; o.mx = L.mx; o.my = L.my; o.mz = L.mz; o.n = L.n;
    o.d = o.db = o.sx = o.sy = o.sz = 0.0;
    o.g = g27;
    o.nblk = 1;
    o.blk[0] = 0;
    return o;
}
static int ilu27_pivots(const LevelInfo& L, const double* g27, double* u, hipStream_t st) {
    const SorOp o = sor_op27(L, g27);
    const int np = (L.mx - 1) + (L.dim >= 2 ? L.my - 1 : 0) + (L.dim == 3 ? L.mz - 1 : 0) + 1;
    const long nt = L.dim == 3 ? (long)L.my * L.mz : (L.dim == 2 ? (long)L.my : 1L);
    const int gb = (int)((nt + 255) / 256);
    for (int p = 0; p < np; ++p) k_ilu27_pivot_plane<<<gb, 256, 0, st>>>(o, u, p);
    LAUNCH_CHECK();
    return 0;
}
template <typename T>
static int ilu27_run(const LevelInfo& L, const double* g27, const double* u, const T* r, T* z, hipStream_t st) {
    if (!g27 || !u) return set_error(PETSC_ERR_LIB, "internal: ILU(0) of a 27-point level without data");
    const SorOp o = sor_op27(L, g27);
    const int np = (L.mx - 1) + (L.dim >= 2 ? L.my - 1 : 0) + (L.dim == 3 ? L.mz - 1 : 0) + 1;
    const long nt = L.dim == 3 ? (long)L.my * L.mz : (L.dim == 2 ? (long)L.my : 1L);
    const int gb = (int)((nt + 255) / 256);
    for (int p = 0; p < np; ++p) k_ilu27_plane<T><<<gb, 256, 0, st>>>(o, u, p, 0, r, z);
    for (int p = np - 1; p >= 0; --p) k_ilu27_plane<T><<<gb, 256, 0, st>>>(o, u, p, 1, r, z);
    LAUNCH_CHECK();
    return 0;
}

// z = (L U)^-1 r with the ILU(0) factor of a constant-operator level
template <typename T>
static int ilu_run(const LevelInfo& L, const double* u, const T* r, T* z, hipStream_t st) {
    if (L.var || L.gal || !u) return set_error(PETSC_ERR_SUP, "ILU(0) is implemented for constant 3/5/7-point levels");
    SorOp o;
    o.dim = L.dim; o.mx = L.mx; o.my = L.my; o.mz = L.mz; o.n = L.n;
    o.d = L.d; o.db = L.db; o.sx = L.sx; o.sy = L.sy; o.sz = L.sz;
    o.g = nullptr;
    o.nblk = 1;
    o.blk[0] = 0;
    const int np = (L.mx - 1) + (L.dim >= 2 ? L.my - 1 : 0) + (L.dim == 3 ? L.mz - 1 : 0) + 1;
    const long nt = L.dim == 3 ? (long)L.my * L.mz : (L.dim == 2 ? (long)L.my : 1L);
    const int gb = (int)((nt + 255) / 256);
    for (int p = 0; p < np; ++p) k_ilu_plane<T><<<gb, 256, 0, st>>>(o, u, p, 0, r, z);
    for (int p = np - 1; p >= 0; --p) k_ilu_plane<T><<<gb, 256, 0, st>>>(o, u, p, 1, r, z);
    LAUNCH_CHECK();
    return 0;
}

// ILU(0) of a 2D box-pattern operator (k_ilu9_factor_plane): factor f of 9n
static int ilu9_factor(const LevelInfo& L, const double* g27, double* f, hipStream_t st) {
    if (L.dim != 2 || !g27 || !f) return set_error(PETSC_ERR_SUP, "box-pattern ILU(0) is implemented in 2D");
    const int np = (L.mx - 1) + 2 * (L.my - 1) + 1, gb = (L.my + 255) / 256;
    for (int p = 0; p < np; ++p) k_ilu9_factor_plane<<<gb, 256, 0, st>>>(L.mx, L.my, g27, f, p);
    LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int ilu9_run(const LevelInfo& L, const double* f, const T* r, T* z, hipStream_t st) {
    if (L.dim != 2 || !f) return set_error(PETSC_ERR_SUP, "box-pattern ILU(0) is implemented in 2D");
    const int np = (L.mx - 1) + 2 * (L.my - 1) + 1, gb = (L.my + 255) / 256;
    for (int p = 0; p < np; ++p) k_ilu9_plane<T><<<gb, 256, 0, st>>>(L.mx, L.my, f, p, 0, r, z);
    for (int p = np - 1; p >= 0; --p) k_ilu9_plane<T><<<gb, 256, 0, st>>>(L.mx, L.my, f, p, 1, r, z);
    LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------- timing

struct KStat {
    long launches = 0, fine_launches = 0;
    double ms = 0, fine_ms = 0, bytes = 0, fine_bytes = 0;
};

struct Timing {
    int on[P4G_K_COUNT] = {};   // 0 off, 1 every launch, 2 finest-level launches only
    KStat st[P4G_K_COUNT];
    struct Pending { hipEvent_t a, b; int cls; bool fine; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    void collect() {
        for (auto& p : pending) {
            float ms = 0.f;
            hipEventSynchronize(p.b);
            hipEventElapsedTime(&ms, p.a, p.b);
            st[p.cls].ms += ms;
            if (p.fine) st[p.cls].fine_ms += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~Timing() {
        for (auto& p : pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
        for (auto e : pool) hipEventDestroy(e);
    }
};

// shared device context of one solver: stream, partials, scalars, timing
struct Ctx {
    hipStream_t stream = nullptr;
    double* part = nullptr;       // block partial sums (2 per block)
    double* part_rr = nullptr;    // second buffer: (r,r) of a fused residual update
    long npart_cap = 0;
    double* S = nullptr;          // scalar block (32 doubles: engine 0..15, outer 16..31)
    double* hS = nullptr;         // pinned mirror
    Timing tm;

    template <class F>
    int launch(int cls, bool fine, double bytes, F&& f) {
        KStat& s = tm.st[cls];
        s.launches++;
        s.bytes += bytes;
        if (fine) { s.fine_launches++; s.fine_bytes += bytes; }
        if (tm.on[cls] == 1 || (tm.on[cls] == 2 && fine)) {
            hipEvent_t a = tm.get(), b = tm.get();
            hipEventRecord(a, stream);
            f();
            hipEventRecord(b, stream);
            tm.pending.push_back({a, b, cls, fine});
        } else {
            f();
        }
        LAUNCH_CHECK();
        return 0;
    }
    int finalize(long nparts, double* Sblk, int mode, int first, const double* src = nullptr) {
        const double* pp = src ? src : part;
        return launch(P4G_K_REDUCE, false, 16.0 * nparts, [&] {
            k_finalize<<<1, FT, 0, stream>>>(nparts, pp, Sblk, mode, first);
        });
    }
    int readS(double* host) {
        HIP_TRY(hipMemcpyAsync(host, S, 32 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return 0;
    }
    // copy the scalar block to host and mark the point; work queued after
    // readS_begin keeps the device busy while the host waits in readS_end
    hipEvent_t sev = nullptr;
    int readS_begin(double* host) {
        if (!sev) HIP_TRY(hipEventCreateWithFlags(&sev, hipEventDisableTiming));
        HIP_TRY(hipMemcpyAsync(host, S, 32 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipEventRecord(sev, stream));
        return 0;
    }
    int readS_end() {
        HIP_TRY(hipEventSynchronize(sev));
        return 0;
    }
};

static inline int vblocks(long n) { return (int)std::max(1L, std::min<long>(VBLOCKS, (n + 255) / 256)); }

// launch geometry of the marching stencil kernels (2D/3D: V x points per lane)
static dim3 stencil_grid_k(int dim, int mx, int my, int mz, int V, int kch) {
    if (dim == 1) return dim3((mx - 1 + TX - 1) / TX, 1, 1);
    const int gx = (mx - 1 + V * TX - 1) / (V * TX);
    if (dim == 3) return dim3(gx, (my + TY - 1) / TY, (mz + kch - 1) / kch);
    return dim3(gx, ((my + kch - 1) / kch + TY - 1) / TY, 1);
}
static dim3 stencil_grid_v(int dim, int mx, int my, int mz, int V) {
    return stencil_grid_k(dim, mx, my, mz, V, kchf(dim, my, mz));
}
// marching chunk of the fused prolongation + smoothing kernel: its per-chunk
// set-up (coarse planes, the first iterates, the halo rows) is longer, so
// chunks are longer than those of the stencil passes (3D, 6-row blocks:
// 257^3 120 us at 16 planes, 134 us at 8; 513^3 855 us at 16, 939 us at 8)
static inline int kch_fused(int dim, int my, int mz) {
    const int slow = dim == 3 ? mz : my;
    return slow >= 200 ? 16 : std::min(16, 2 * kchf(dim, my, mz));
}
template <typename T>
static dim3 stencil_grid(int dim, int mx, int my, int mz) { return stencil_grid_v(dim, mx, my, mz, VecOf<T>::V); }
// two points per lane: the largest stencil grid
static dim3 pair_grid(int dim, int mx, int my, int mz) { return stencil_grid_v(dim, mx, my, mz, 2); }
// the fused prolongation + smoothing kernel (two points per lane)
static dim3 fused_grid(int dim, int mx, int my, int mz) {
    const int kch = kch_fused(dim, my, mz);
    if (dim == 3) return dim3((mx - 1 + 2 * TX - 1) / (2 * TX), (my + TYF - 1) / TYF, (mz + kch - 1) / kch);
    return stencil_grid_k(dim, mx, my, mz, 2, kch);
}
static dim3 stencil_block(int dim) { return dim == 1 ? dim3(TX, 1, 1) : dim3(TX, TY, 1); }
static long nblocks(dim3 g) { return (long)g.x * g.y * g.z; }
// the most blocks of any pass that writes block partial sums, over the
// levels (the marching chunks depend on the size, so not always the finest)
static long max_part_blocks(const std::vector<LevelInfo>& lv) {
    long m = 0;
    for (const LevelInfo& L : lv)
        m = std::max(m, std::max(nblocks(pair_grid(L.dim, L.mx, L.my, L.mz)), nblocks(fused_grid(L.dim, L.mx, L.my, L.mz))));
    return m;
}
// transfers: lanes over coarse columns I = 1..mxc-1, rows and march chunks
// over the grid (my, mz) the kernel walks (coarse: restriction, fine: prolongation)
static dim3 transfer_grid(int dim, int mxc, int my, int mz, int kch) {
    const int gx = std::max(1, (mxc - 1 + TX - 1) / TX);
    if (dim == 3) return dim3(gx, (my + TY - 1) / TY, (mz + kch - 1) / kch);
    if (dim == 2) return dim3(gx, ((my + kch - 1) / kch + TY - 1) / TY, 1);
    return dim3(gx, 1, 1);
}
// marching chunk of the restriction, in coarse planes (rows in 2D) of the
// coarse grid (my, mz): each is two fine planes read (fine 257^3: 45 us at
// 1-2, 48 us at 4; fine 513^3: 293 us at 4, 303 us at 2 or 8)
static inline int kch_restrict(int dim, int my, int mz) {
    const int slow = dim == 3 ? mz : (dim == 2 ? my : 1);
    return slow >= 200 ? 4 : (slow >= 100 ? 2 : 1);
}

// x += alpha_0 p_0 + ... + alpha_{m-1} p_{m-1} (f.m = m terms, in order:
// deferred CG iterate updates; init: x = the sum, x not read), then the
// list is empty
template <typename P>
static int launch_fold_x(Ctx* cx, long n, double* x, FoldListT<P>& f, bool init = false) {
    hipStream_t st = cx->stream;
    const int gb = vblocks((n + 3) / 4);
    const int in = init ? 1 : 0;
    if (f.m < 1 || f.m > kFoldMax) return set_error(PETSC_ERR_LIB, "internal: fold list of %d", f.m);
    P4G_TRY(cx->launch(P4G_K_OTHER, true, ((init ? 8.0 : 16.0) + sizeof(P) * f.m) * n, [&] {
        switch (f.m) {
        case 1: k_fold_x<1, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 2: k_fold_x<2, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 3: k_fold_x<3, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 4: k_fold_x<4, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 5: k_fold_x<5, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 6: k_fold_x<6, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 7: k_fold_x<7, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 8: k_fold_x<8, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 9: k_fold_x<9, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 10: k_fold_x<10, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 11: k_fold_x<11, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 12: k_fold_x<12, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 13: k_fold_x<13, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 14: k_fold_x<14, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        case 15: k_fold_x<15, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        default: k_fold_x<16, P><<<gb, 256, 0, st>>>(n, x, f, in); break;
        }
    }));
    f.m = 0;
    return 0;
}

// ------------------------------------------------------------- engine

// Device hierarchy and Krylov work vectors in precision T.
template <typename T>
struct Engine {
    Ctx* cx = nullptr;
    std::vector<LevelInfo>* lv = nullptr;
    int nlev = 0, cycles = 1, smoother = SM_CHEB, nu = 1, jacobi = 1, pctype = PCT_MG;
    double rich_scale = 1.0;
    bool diff_form = false;   // constant-coefficient rows in difference form
    // V(1,1): the post-smoothing step fused with the prolongation (with more
    // steps the prolonged iterate would have to be stored for the Chebyshev
    // recurrence, and the fused kernel then loses to the separate passes)
    bool fuse_post = true;
    // CG updates fused into stencil passes: the residual update into the
    // first kernel of the cycle (fuse_r), the direction update with the
    // operator apply (fuse_dir)
    bool fuse_r = true, fuse_dir = true;
    // 3D constant-coefficient passes with the source staged in LDS
    // (k_stencil_lds): -1 on levels of 400 or more points per side, 1 on
    // every level whose rows fill the waves, 0 never (-p4g_lds)
    int lds = -1;
    // the residual and restriction of a cycle in one pass (-p4g_resid_restrict),
    // with rr_rows coarse rows per block (-p4g_rr_rows 8|4)
    int resid_restrict_mode = -1;
    int rr_rows = 8;
    int rr_planes = 0;   // its coarse planes per block (-p4g_rr_planes; 0: by size)
    // the prolongation and the first of several post-smoothing steps in one
    // pass (-p4g_post_fuse)
    int post2_mode = -1;
    // per level: 27-point Galerkin coefficients (FP64, device) or null
    std::vector<const double*> g27;
    bool fmg = false;          // full multigrid: an F-cycle per apply
    struct Dev {
        T *x = nullptr, *b = nullptr, *r = nullptr, *t[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        T* kc = nullptr;
    };
    std::vector<Dev> dv;
    double* Ainv = nullptr;   // dense inverse of the coarsest operator
    int n0 = 0;
    // reduced-space VI step: diagonal of the finest (null-row) operator,
    // zero at active vertices; null otherwise
    const double* vidiag = nullptr;
    const double* ilu9 = nullptr;   // a 2D box-pattern ILU(0) factor of the finest level, when set
    // a matrix-free operator on the finest level (-snes_mf_operator): y = A u
    // formed by mf_op into mf_tmp, the pass's output applied to it; the
    // level's 27-plane table is then only the preconditioning matrix
    std::function<int(const T*, T*)> mf_op;
    T* mf_tmp = nullptr;
    std::function<int(const T*, T*)> asm_op;   // PCASM's z = B r (set by the VI solve)
    // CG (FP64): the iterate updates x += alpha p of up to x_defer iterations
    // deferred and applied in order (directions kept in pdef)
    int x_defer = 0;
    T* pdef[kFoldMax] = {};
    bool mf_bypass = false;   // a pass on the preconditioning matrix (PCApplyRichardson_SOR's sweeps)
    bool vi = false;   // levels may become null-row Galerkin levels later (full temporaries)
    int mcycles = 1;        // -pc_mg_multiplicative_cycles
    bool kaskade = false;   // -pc_mg_type kaskade
    bool full_temps = false;   // four smoother temporaries on every level
    bool sor = false;          // explicit smoother path: SSOR or ILU(0) level preconditioner
    bool lilu = false;         // the level preconditioner is ILU(0) (-mg_levels_pc_type ilu|icc)
    int coarse_its = 0;        // coarsest level: >0 PCG iterations (-mg_coarse_ksp_type cg), else the dense solve
    int coarse_pc = 0;         //   its PC: 0 none, 1 Jacobi, 2 ILU(0)
    std::vector<const double*> ilu_u;   // per level: ILU(0) pivots (device) or null
    // Krylov vectors on the finest level
    T *r = nullptr, *z = nullptr, *pa = nullptr, *pb = nullptr, *w = nullptr, *xbuf = nullptr,
      *bbuf = nullptr, *r2 = nullptr;
    // A CG residual update r_new = r - alpha w pending for the next
    // preconditioner application: the first fine-level kernel that reads the
    // residual forms r_new on the fly and stores it (rr: also (r_new, r_new)
    // into cx->part_rr, nparts blocks).
    struct Pend {
        const T* r = nullptr;
        const T* w = nullptr;
        const double* alpha = nullptr;
        T* rn = nullptr;
        bool rr = false, on = false;
        long nparts = 0;
    } pend;
    std::vector<void*> allocs;

    ~Engine() { for (void* p : allocs) hipFree(p); }

    int alloc(T** p, long n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<long>(n, 1) * sizeof(T));
        if (e != hipSuccess)
            return set_error(PETSC_ERR_MEM, "hipMalloc of %ld bytes failed: %s", n * (long)sizeof(T),
                             hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return 0;
    }

    bool zg_fusable(const LevelInfo& L) const {
        return !sor && !L.var && !L.gal && L.d == L.db && (smoother == SM_CHEB || smoother == SM_RICH);
    }

    int setup(Ctx* c, std::vector<LevelInfo>* levels, const std::vector<double>& Ainv_host,
              int n0_) {
        cx = c;
        lv = levels;
        nlev = (int)levels->size();
        dv.resize(nlev);
        const LevelInfo& F = (*lv)[nlev - 1];
        for (int l = 0; l < nlev; ++l) {
            const LevelInfo& L = (*lv)[l];
            Dev& d = dv[l];
            if (pctype == PCT_MG) {
                if (l < nlev - 1) { P4G_TRY(alloc(&d.x, L.n)); P4G_TRY(alloc(&d.b, L.n)); }
                if (l > 0 || nlev == 1 || coarse_its > 0) {   // smoothing levels, an iterative coarse solve
                    P4G_TRY(alloc(&d.r, L.n));
                    int nt = sor ? 5 : (!vi && !full_temps && nu == 1 && cycles == 1 && zg_fusable(L)) ? 1 : 4;
                    if (l == 0 && nlev > 1) nt = 3;   // the coarse CG's z, p, w
                    for (int q = 0; q < nt; ++q) P4G_TRY(alloc(&d.t[q], L.n));
                }
            }
            if (L.var) {
                const long ncell = (long)L.cx * L.cy * L.cz;
                P4G_TRY(alloc(&d.kc, ncell));
                std::vector<T> tmp(L.kcell.begin(), L.kcell.end());
                HIP_TRY(hipMemcpy(d.kc, tmp.data(), ncell * sizeof(T), hipMemcpyHostToDevice));
            }
        }
        P4G_TRY(alloc(&r, F.n));
        P4G_TRY(alloc(&z, F.n));
        P4G_TRY(alloc(&pa, F.n));
        P4G_TRY(alloc(&pb, F.n));
        P4G_TRY(alloc(&w, F.n));
        P4G_TRY(alloc(&xbuf, F.n));
        P4G_TRY(alloc(&r2, F.n));
        if (x_defer < 0 || x_defer > kFoldMax - 1) return set_error(PETSC_ERR_LIB, "internal: x_defer %d", x_defer);
        for (int j = 0; j < x_defer; ++j) P4G_TRY(alloc(&pdef[j], F.n));
        if (pctype == PCT_MG) {
            n0 = n0_;
            void* q = nullptr;
            HIP_TRY(hipMalloc(&q, Ainv_host.size() * sizeof(double)));
            allocs.push_back(q);
            Ainv = (double*)q;
            HIP_TRY(hipMemcpy(Ainv, Ainv_host.data(), Ainv_host.size() * sizeof(double),
                              hipMemcpyHostToDevice));
        }
        return 0;
    }

    // a new coarsest operator of the same size (reduced-space VI steps)
    int set_coarse_inverse(const std::vector<double>& Ainv_host) {
        if (!Ainv || (long)Ainv_host.size() != (long)n0 * n0)
            return set_error(PETSC_ERR_LIB, "internal: coarse inverse of a different size");
        HIP_TRY(hipMemcpyAsync(Ainv, Ainv_host.data(), Ainv_host.size() * sizeof(double), hipMemcpyHostToDevice,
                               cx->stream));
        HIP_TRY(hipStreamSynchronize(cx->stream));
        return 0;
    }

    OpConst<T> opc(const LevelInfo& L, bool jac) const {
        OpConst<T> o;
        o.d = T(L.d); o.db = T(L.db); o.sx = T(L.sx); o.sy = T(L.sy); o.sz = T(L.sz);
        o.id = jac ? T(1.0 / L.d) : T(1); o.idb = jac ? T(1.0 / L.db) : T(1);
        // difference form only where it is the same row to working precision
        o.diff = diff_form && std::fabs(L.d - 2.0 * (L.sx + L.sy + L.sz)) <= 1e-12 * std::fabs(L.d) ? 1 : 0;
        return o;
    }
    OpVar<T> opv(int l, bool jac) const {
        const LevelInfo& L = (*lv)[l];
        OpVar<T> o;
        o.kc = dv[l].kc; o.cx = L.cx; o.cy = L.cy; o.cz = L.cz;
        o.ax = T(L.ax); o.ay = T(L.ay); o.az = T(L.az); o.jac = jac ? 1 : 0;
        return o;
    }

    // whether a 3D constant-coefficient pass on level L runs k_stencil_lds:
    // the rows fill whole waves ((mx - 1) a multiple of 64 V) and -p4g_lds
    // asks for it (auto: levels of 400 or more points per side, where the
    // operand misses the last-level cache between passes)
    bool lds_level(const LevelInfo& L) const {
        constexpr int W = TX * VecOf<T>::V;
        if (L.dim != 3 || L.var || L.gal || lds == 0 || (L.mx - 1) % W != 0) return false;
        return lds == 1 || L.mz >= 400;
    }

    // a pass of the finest level under a matrix-free operator: the product,
    // then the output per point (plain sources only: the operand must exist)
    template <class Src, class Out>
    int mf_stencil(Src src, Out out, int cls, double passes, long* nb, bool jac, double* pbuf) {
        const LevelInfo& L = (*lv)[nlev - 1];
        const T* u = nullptr;
        const T* au = nullptr;
        if constexpr (std::is_same<Src, SrcArr<T>>::value) {
            P4G_TRY(mf_op(src.x, mf_tmp));
            u = src.x;
            au = mf_tmp;
        } else if constexpr (!std::is_same<Src, SrcZero<T>>::value) {
            return set_error(PETSC_ERR_LIB, "internal: a fused source over the matrix-free operator");
        }
        if (!mf_tmp || g27.size() < (size_t)nlev || !g27[nlev - 1])
            return set_error(PETSC_ERR_LIB, "internal: matrix-free level without its buffers");
        const int gb = vblocks(L.n);
        if (nb) *nb = gb;
        double* part = pbuf ? pbuf : cx->part;
        hipStream_t s = cx->stream;
        const double* gp = g27[nlev - 1];
        return cx->launch(cls, true, passes * L.n * sizeof(T), [&] {
            k_out_au<T, Out><<<gb, 256, 0, s>>>(L.n, u, au, gp, jac ? 1 : 0, out, part);
        });
    }

    // one marching stencil pass on level l
    template <class Src, class Out>
    int stencil(int l, Src src, Out out, int cls, double passes, long* nb, bool jac, double* pbuf = nullptr) {
        const LevelInfo& L = (*lv)[l];
        if (mf_op && !mf_bypass && l == nlev - 1) return mf_stencil(src, out, cls, passes, nb, jac, pbuf);
        if (L.gal) {   // a Galerkin level: the 27-point pass
            const int gb = vblocks(L.n);
            if (nb) *nb = gb;
            const Op27<T> o{g27[l], L.n, L.mx, L.my, L.mz, jac ? 1 : 0};
            hipStream_t s = cx->stream;
            double* part = pbuf ? pbuf : cx->part;
            const double bytes = passes * L.n * sizeof(T) + 27.0 * L.n * sizeof(double);
            return cx->launch(cls, l == nlev - 1, bytes,
                              [&] { k_stencil27<T, Src, Out><<<gb, 256, 0, s>>>(o, src, out, part); });
        }
        const dim3 g = stencil_grid<T>(L.dim, L.mx, L.my, L.mz), b = stencil_block(L.dim);
        if (nb) *nb = nblocks(g);
        if (Out::NDOT && nblocks(g) > cx->npart_cap)
            return set_error(PETSC_ERR_LIB, "internal: partial-sum buffer too small");
        hipStream_t s = cx->stream;
        double* part = pbuf ? pbuf : cx->part;
        const bool fine = (l == nlev - 1);
        const double bytes = passes * L.n * sizeof(T);
        if (L.var) {
            if (L.dim != 3) return set_error(PETSC_ERR_SUP, "variable coefficient needs 3D");
            OpVar<T> o = opv(l, jac);
            return cx->launch(cls, fine, bytes, [&] {
                k_stencil<T, 3, OpVar<T>, Src, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), o, src, out,
                                                                 part);
            });
        }
        OpConst<T> o = opc(L, jac);
        if constexpr (lds_stageable<Src>::value) {
            if (lds_level(L)) {   // the same grid and block as k_stencil (TYD = TY)
                return cx->launch(cls, fine, bytes, [&] {
                    k_stencil_lds<T, OpConst<T>, Src, Out, TY, 1><<<g, b, 0, s>>>(
                        L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), o, src, out, part);
                });
            }
        }
        return cx->launch(cls, fine, bytes, [&] {
            if (L.dim == 3)
                k_stencil<T, 3, OpConst<T>, Src, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), o, src,
                                                                   out, part);
            else if (L.dim == 2)
                k_stencil<T, 2, OpConst<T>, Src, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), o, src,
                                                                   out, part);
            else
                k_stencil<T, 1, OpConst<T>, Src, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), o, src,
                                                                   out, part);
        });
    }

    // the fused residual + restriction (k_resid_restrict_lds) on level l:
    // 3D constant coefficient, rows of 64 V fine points, -p4g_resid_restrict
    // (auto: where the LDS-staged passes run)
    bool rr_fusable(const LevelInfo& L) const {
        if (L.dim != 3 || L.var || L.gal || resid_restrict_mode == 0 || (L.mx - 1) % (TX * VecOf<T>::V) != 0)
            return false;
        return resid_restrict_mode == 1 || lds_level(L);
    }

    // b_c = P^T (b - A x), level l -> l-1, the fine residual not stored
    // (rr_rows coarse rows per block: 8, or 4)
    int resid_restrict(int l, const T* xx, const T* bb, T* bc) {
        const LevelInfo& F = (*lv)[l];
        const LevelInfo& C = (*lv)[l - 1];
        const int tyc = rr_rows == 4 ? 4 : 8;
        const int kch = rr_planes > 0 ? rr_planes : kch_restrict(C.dim, C.my, C.mz);
        constexpr int V = VecOf<T>::V;   // fine points per lane (16-byte accesses)
        if ((F.mx - 1) % (V * TX) != 0) return set_error(PETSC_ERR_LIB, "internal: residual + restriction tile width");
        const dim3 g((F.mx - 1) / (V * TX), (C.my + tyc - 1) / tyc, (C.mz + kch - 1) / kch), b(TX, tyc);
        hipStream_t s = cx->stream;
        OpConst<T> o = opc(F, false);
        return cx->launch(P4G_K_RESID_ZG, l == nlev - 1, double(2 * F.n + C.n) * sizeof(T), [&] {
            if (tyc == 4)
                k_resid_restrict_lds<T, V, OpConst<T>, 4><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, C.mz, kch, o,
                                                                         xx, bb, bc);
            else
                k_resid_restrict_lds<T, V, OpConst<T>, 8><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, C.mz, kch, o,
                                                                         xx, bb, bc);
        });
    }

    int restrict_(int l, const T* rf, T* bc) {   // level l -> l-1
        const LevelInfo& F = (*lv)[l];
        const LevelInfo& C = (*lv)[l - 1];
        const int kch = kch_restrict(C.dim, C.my, C.mz);
        const dim3 g = transfer_grid(C.dim, C.mx, C.my, C.mz, kch), b = stencil_block(C.dim);
        hipStream_t s = cx->stream;
        return cx->launch(P4G_K_RESTRICT, l == nlev - 1, double(F.n + C.n) * sizeof(T), [&] {
            if (F.dim == 3)
                k_restrict<T, 3><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, C.mz, kch, rf, bc);
            else if (F.dim == 2)
                k_restrict<T, 2><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, C.mz, kch, rf, bc);
            else
                k_restrict<T, 1><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, C.mz, kch, rf, bc);
        });
    }

    // One smoothing step (Out) applied to t = (zg ? c base : base) + P x_c
    // on level l without storing t.
    template <bool ZG, class Out>
    int prolong_smooth(int l, const T* base, T cz, const T* xc, Out out, double passes, long* nb, bool jac) {
        const LevelInfo& L = (*lv)[l];
        const LevelInfo& C = (*lv)[l - 1];
        const dim3 g = fused_grid(L.dim, L.mx, L.my, L.mz), b = stencil_block(L.dim);
        if (nb) *nb = nblocks(g);
        if (Out::NDOT && nblocks(g) > cx->npart_cap)
            return set_error(PETSC_ERR_LIB, "internal: partial-sum buffer too small");
        hipStream_t s = cx->stream;
        const double bytes = passes * L.n * sizeof(T) + C.n * sizeof(T);
        const int kch = kch_fused(L.dim, L.my, L.mz);
        if (L.var) {
            if (L.dim != 3) return set_error(PETSC_ERR_SUP, "variable coefficient needs 3D");
            OpVar<T> o = opv(l, jac);
            return cx->launch(P4G_K_PROLONG, l == nlev - 1, bytes, [&] {
                k_prolong_smooth<T, 3, OpVar<T>, ZG, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, C.mx, C.my, kch, o, base,
                                                                         cz, xc, out, cx->part);
            });
        }
        OpConst<T> o = opc(L, jac);
        constexpr int V = VecOf<T>::V;   // fine points per lane (16-byte accesses)
        if (lds_level(L) && (L.mx - 1) % (V * TX) == 0) {   // LDS-staged form: blocks of TY output rows
            const dim3 gl((L.mx - 1) / (V * TX), (L.my + TY - 1) / TY, (L.mz + kch - 1) / kch), bl(TX, TY);
            if (nb) *nb = nblocks(gl);
            if (Out::NDOT && nblocks(gl) > cx->npart_cap)
                return set_error(PETSC_ERR_LIB, "internal: partial-sum buffer too small");
            return cx->launch(P4G_K_PROLONG, l == nlev - 1, bytes, [&] {
                k_prolong_smooth_lds<T, V, OpConst<T>, ZG, Out, TY><<<gl, bl, 0, s>>>(
                    L.mx, L.my, L.mz, C.mx, C.my, kch, o, base, cz, xc, out, cx->part);
            });
        }
        return cx->launch(P4G_K_PROLONG, l == nlev - 1, bytes, [&] {
            if (L.dim == 3)
                k_prolong_smooth<T, 3, OpConst<T>, ZG, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, C.mx, C.my, kch, o, base,
                                                                           cz, xc, out, cx->part);
            else
                k_prolong_smooth<T, 2, OpConst<T>, ZG, Out><<<g, b, 0, s>>>(L.mx, L.my, L.mz, C.mx, C.my, kch, o, base,
                                                                           cz, xc, out, cx->part);
        });
    }

    // the first post-smoothing step fused with the prolongation (2D/3D):
    // y = t + s D^-1 (b - A t), t = pre + P x_c, t not stored
    template <bool DOTS>
    int post_first(int l, bool zg, const T* pre, T cz, const T* xc, const T* b, T* y, T s2, long* nb) {
        const bool jac = jacobi != 0;
        typedef OutCheb<T, false, DOTS ? 2 : 0> OC;
        const OC oc{b, nullptr, y, T(0), T(1), s2};
        const double passes = zg ? 2.0 : 3.0;
        return zg ? prolong_smooth<true>(l, pre, cz, xc, oc, passes, nb, jac)
                  : prolong_smooth<false>(l, pre, cz, xc, oc, passes, nb, jac);
    }

    // out = (zg ? c b : base) + P x_c   (level l-1 -> l); on the finest level
    // of a reduced-space VI step the active vertices of out are zeroed
    int prolong(int l, bool zg, const T* base, T c, const T* xc, T* out) {
        P4G_TRY(prolong_q1(l, zg, base, c, xc, out));
        if (l != nlev - 1 || !vidiag) return 0;
        const long n = (*lv)[l].n;
        hipStream_t s = cx->stream;
        return cx->launch(P4G_K_PROLONG, true, double(n) * (8.0 + sizeof(T)), [&] {
            k_vi_mask<T><<<vblocks(n), 256, 0, s>>>(n, vidiag, out);
        });
    }

    int prolong_q1(int l, bool zg, const T* base, T c, const T* xc, T* out) {
        const LevelInfo& F = (*lv)[l];
        const LevelInfo& C = (*lv)[l - 1];
        const dim3 g = transfer_grid(F.dim, C.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz)), b = stencil_block(F.dim);
        hipStream_t s = cx->stream;
        return cx->launch(P4G_K_PROLONG, l == nlev - 1, double(2 * F.n + C.n) * sizeof(T), [&] {
            if (F.dim == 3) {
                if (zg) k_prolong<T, 3, true><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
                else k_prolong<T, 3, false><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
            } else if (F.dim == 2) {
                if (zg) k_prolong<T, 2, true><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
                else k_prolong<T, 2, false><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
            } else {
                if (zg) k_prolong<T, 1, true><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
                else k_prolong<T, 1, false><<<g, b, 0, s>>>(F.mx, F.my, F.mz, C.mx, C.my, kchf(F.dim, F.my, F.mz), base, c, xc, out);
            }
        });
    }

    // the coarsest level's KSP CG with a fixed iteration count (convergence
    // test skip), from x (xzero: from 0): the oracle's KSPSolve_CG without norms
    int coarse_cg(const T* b, T* x, bool xzero) {
        const LevelInfo& L = (*lv)[0];
        const long n = L.n;
        Dev& d = dv[0];
        T *r = d.r, *z = d.t[0], *p = d.t[1], *w = d.t[2];
        if (!r || !z || !p || !w) return set_error(PETSC_ERR_LIB, "internal: coarse CG vectors");
        hipStream_t st = cx->stream;
        const int gb = vblocks(n);
        double* Sb = cx->S + 16;            // the second scalar block (the outer KSP uses the first)
        const double* hs = cx->hS + 16;
        if (xzero) {
            P4G_TRY(fill(n, T(0), x));
            HIP_TRY(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        } else {
            P4G_TRY(stencil(0, SrcArr<T>{x}, OutResid<T, 0>{b, r}, P4G_K_COARSE, 3, nullptr, false));
        }
        auto zr = [&](double& beta) -> int {   // z = B r, beta = (z, r)
            if (coarse_pc == 2) P4G_TRY(ilu_run<T>(L, ilu_u.size() > 0 ? ilu_u[0] : nullptr, r, z, st));
            else if (coarse_pc == 1) k_jacobi_const<T><<<gb, 256, 0, st>>>(L.dim, L.mx, L.my, L.mz, L.d, L.db, r, z);
            else HIP_TRY(hipMemcpyAsync(z, r, n * sizeof(T), hipMemcpyDeviceToDevice, st));
            k_dot2<T><<<gb, 256, 0, st>>>(n, z, r, cx->part);
            LAUNCH_CHECK();
            P4G_TRY(cx->finalize(gb, Sb, FIN_RAW, 0));
            P4G_TRY(cx->readS(cx->hS));
            beta = hs[S_TMP0];
            return 0;
        };
        double beta = 0.0, betaold = 0.0, dpi = 0.0, dpiold = 0.0;
        P4G_TRY(zr(beta));
        for (int i = 0; i < coarse_its; ++i) {
            if (beta == 0.0) break;
            const double bb = i == 0 ? 0.0 : beta / betaold;
            k_comb3<T><<<gb, 256, 0, st>>>(n, 0.0, nullptr, bb, i == 0 ? nullptr : p, 1.0, z, p);
            long nbs = 0;
            P4G_TRY(stencil(0, SrcArr<T>{p}, OutApplyDot<T>{w}, P4G_K_COARSE, 2, &nbs, false));
            P4G_TRY(cx->finalize(nbs, Sb, FIN_RAW, 0));
            P4G_TRY(cx->readS(cx->hS));
            dpiold = dpi;
            dpi = hs[S_TMP0];
            betaold = beta;
            if (dpi == 0.0 || (i > 0 && ((dpi > 0) != (dpiold > 0)))) break;
            const double a = beta / dpi;
            k_axpy<T><<<gb, 256, 0, st>>>(n, T(a), p, x);
            LAUNCH_CHECK();
            if (i + 1 == coarse_its) break;   // the last B r would go unused
            k_axpy<T><<<gb, 256, 0, st>>>(n, T(-a), w, r);
            LAUNCH_CHECK();
            P4G_TRY(zr(beta));
        }
        return 0;
    }

    int coarse(const T* b, T* x, bool xzero = true) {
        if (coarse_its > 0) return coarse_cg(b, x, xzero);
        hipStream_t s = cx->stream;
        return cx->launch(P4G_K_COARSE, nlev == 1, double(n0) * n0 * 8.0, [&] {
            k_coarse<T><<<(n0 + 3) / 4, 256, 0, s>>>(n0, Ainv, b, x);
        });
    }

    int scale(long n, T c, const T* x, T* y, bool dots, bool fine) {
        hipStream_t s = cx->stream;
        const int gb = vblocks(n);
        return cx->launch(dots ? P4G_K_SMOOTH : P4G_K_OTHER, fine, 2.0 * n * sizeof(T), [&] {
            if (dots) k_scale<T, 2><<<gb, 256, 0, s>>>(n, c, x, y, cx->part);
            else k_scale<T, 0><<<gb, 256, 0, s>>>(n, c, x, y, cx->part);
        });
    }

    int fill(long n, T v, T* x) {
        hipStream_t s = cx->stream;
        return cx->launch(P4G_K_OTHER, false, double(n) * sizeof(T), [&] {
            k_fill<T><<<vblocks(n), 256, 0, s>>>(n, v, x);
        });
    }

    // Chebyshev / Richardson smoothing with Jacobi on level l:
    //   x0 = nullptr means a zero initial guess.  The result lands in `out`.
    //   DOTS: fuse (out, b) and (out, out) partials into the final step.
    // mstart = 2: continue after a first step done elsewhere; x0 = iterate 1,
    // pm2init = iterate 0.
    // Chebyshev / Richardson smoothing preconditioned by SSOR: each step
    // forms r = b - A p explicitly, z = SSOR(r), then the recurrence (the
    // oracle's KSPSolve_Chebyshev / _Richardson with PCSOR)
    int smooth_sor(int l, const T* b, const T* x0, T* out, bool dots, long* nb) {
        const LevelInfo& L = (*lv)[l];
        const long n = L.n;
        Dev& d = dv[l];
        hipStream_t st = cx->stream;
        T* tt[5];
        int nt = 0;
        for (int q = 0; q < 5; ++q)
            if (d.t[q] && d.t[q] != b && d.t[q] != x0 && d.t[q] != out) tt[nt++] = d.t[q];
        if (nt < 4 || !d.r || d.r == b || d.r == out)
            return set_error(PETSC_ERR_LIB, "internal: SOR smoother temporaries");
        T* r = d.r;
        T* z = tt[0];
        double s, mu = 0.0, omegaprod = 0.0, c0 = 1.0, c1 = 0.0;
        if (smoother == SM_CHEB) {
            s = 2.0 / (L.emax + L.emin);
            const double alpha = 1.0 - s * L.emin;
            mu = 1.0 / alpha;
            omegaprod = 2.0 / alpha;
            c1 = mu;
        } else {
            s = rich_scale;
        }
        const T* pm1 = x0;          // iterate m-1 (nullptr: zero)
        const T* pm2 = nullptr;     // iterate m-2
        const int gb = vblocks(n);
        // Richardson with SOR and scale 1 is PETSc's PCApplyRichardson_SOR:
        // sweeps on the PC's matrix, so a matrix-free operator is not used
        const bool pmat_resid = smoother == SM_RICH && rich_scale == 1.0 && !lilu;
        for (int m = 1; m <= nu; ++m) {
            const T* rr = b;
            if (pm1) {
                mf_bypass = pmat_resid;
                const int e = stencil(l, SrcArr<T>{pm1}, OutResid<T, 0>{b, r}, P4G_K_SMOOTH, 3, nullptr, false);
                mf_bypass = false;
                P4G_TRY(e);
                rr = r;
            }
            if (lilu) P4G_TRY(ilu_run<T>(L, ilu_u.size() > (size_t)l ? ilu_u[l] : nullptr, rr, z, st));
            else P4G_TRY(sor_run<T>(L, g27.size() > (size_t)l ? g27[l] : nullptr, rr, z, st));
            T* dst = out;
            if (m < nu)
                for (int q = 1; q < 4; ++q)
                    if (tt[q] != pm1 && tt[q] != pm2) { dst = tt[q]; break; }
            double ca = 0.0, cb = 1.0, cc = s;
            if (smoother == SM_CHEB && m >= 2) {
                const double ckp1 = 2.0 * mu * c1 - c0;
                const double omega = omegaprod * c1 / ckp1;
                c0 = c1;
                c1 = ckp1;
                ca = 1.0 - omega; cb = omega; cc = omega * s;
            }
            const T* pa = (smoother == SM_CHEB && m >= 2) ? pm2 : nullptr;
            P4G_TRY(cx->launch(P4G_K_SMOOTH, l == nlev - 1, 3.0 * n * sizeof(T), [&] {
                k_comb3<T><<<gb, 256, 0, st>>>(n, ca, pa, cb, pm1, cc, z, dst);
            }));
            pm2 = pm1;
            pm1 = dst;
        }
        if (dots) {
            if (nb) *nb = gb;
            P4G_TRY(cx->launch(P4G_K_SMOOTH, l == nlev - 1, 2.0 * n * sizeof(T), [&] {
                k_dot2<T><<<gb, 256, 0, st>>>(n, out, b, cx->part);
            }));
        }
        return 0;
    }

    int smooth(int l, const T* b, const T* x0, T* out, bool dots, long* nb, int mstart = 1,
               const T* pm2init = nullptr) {
        if (sor) return smooth_sor(l, b, x0, out, dots, nb);
        const LevelInfo& L = (*lv)[l];
        const bool jac = jacobi != 0;
        const double idg = jac ? 1.0 / L.d : 1.0;
        const bool fine = l == nlev - 1;
        double s, alpha = 0, mu = 0, omegaprod = 0, c[3];
        if (smoother == SM_CHEB) {
            s = 2.0 / (L.emax + L.emin);
            alpha = 1.0 - s * L.emin;
            mu = 1.0 / alpha;
            omegaprod = 2.0 / alpha;
            c[0] = 1.0; c[1] = mu;
        } else {
            s = rich_scale;
        }
        // buffers: p[m] holds iterate m; choose temps distinct from inputs
        const T* pm2 = pm2init;          // iterate m-2
        const T* pm1 = x0;               // iterate m-1 (nullptr: zero)
        bool pm1_implicit = (x0 == nullptr);   // iterate 1 of a zero guess kept implicit (c b)
        int used = -1;
        auto pick = [&](const T* a, const T* bq) -> T* {
            for (int q = 0; q < 4; ++q) {
                T* t = dv[l].t[q];
                if (t && t != a && t != bq && t != out) return t;
            }
            return nullptr;
        };
        for (int m = mstart; m <= nu; ++m) {
            const bool last = (m == nu);
            // kernel class: the cycle's last step (with its dot products) apart
            const int scls = (last && dots) ? P4G_K_SMOOTH_LAST : P4G_K_SMOOTH;
            T* dst = last ? out : pick(pm1, pm2);
            if (!dst) return set_error(PETSC_ERR_LIB, "internal: smoother buffers exhausted");
            double c0 = 0.0, c1 = 1.0, c2 = s;
            if (smoother == SM_CHEB && m >= 2) {
                const double ckp1 = 2.0 * mu * c[1] - c[0];
                const double omega = omegaprod * c[1] / ckp1;
                c[0] = c[1]; c[1] = ckp1;
                c0 = 1.0 - omega; c1 = omega; c2 = omega * s;
            }
            const bool prev = (smoother == SM_CHEB && m >= 2 && pm2 != nullptr);
            long nbk = 0;
            if (m == 1 && x0 == nullptr) {
                // iterate 1 from zero: s D^-1 b, pointwise; keep implicit if
                // the next step can form it on the fly
                if (!last && zg_fusable(L) && !(nu >= 3 && smoother == SM_CHEB)) {
                    pm1 = nullptr;
                    pm1_implicit = true;
                    continue;
                }
                if (L.var || L.gal) {
                    // variable diagonal: a stencil pass with u = 0 gives
                    // y = c2 D^-1 b
                    if (last && dots)
                        P4G_TRY(stencil(l, SrcZero<T>{}, OutCheb<T, false, 2>{b, nullptr, dst, T(0), T(0), T(c2)},
                                        scls, 2, &nbk, jac));
                    else
                        P4G_TRY(stencil(l, SrcZero<T>{}, OutCheb<T, false, 0>{b, nullptr, dst, T(0), T(0), T(c2)},
                                        scls, 2, &nbk, jac));
                } else {
                    P4G_TRY(scale(L.n, T(s * idg), b, dst, last && dots, fine));
                    nbk = vblocks(L.n);
                }
                if (nb) *nb = nbk;
                pm2 = nullptr;
                pm1 = dst;
                pm1_implicit = false;
                continue;
            }
            // the m-th step reads iterate m-1 (array or implicit c b) and m-2
            if (pm1_implicit) {
                const T cb = T(s * idg);
                // m == 2 from zero guess: prev iterate is 0, so no prev term
                if (nu >= 3 && smoother == SM_CHEB) return set_error(PETSC_ERR_LIB, "internal smoother path");
                if (fine && pend.on && !(last && dots)) {   // b = r - alpha w, formed here and stored to b
                    if (pend.rn != b) return set_error(PETSC_ERR_LIB, "internal: pending update not into b");
                    const SrcUpd<T> su{pend.r, pend.w, pend.alpha, pend.rn};
                    typedef OutScaleU<T, OutCheb<T, false, 0>> OC;
                    const OC oc{cb, OutCheb<T, false, 0>{b, nullptr, dst, T(c0), T(c1), T(c2)}};
                    pend.on = false;
                    if (pend.rr)
                        P4G_TRY(stencil(l, su, OutRR<T, OC>{oc}, P4G_K_SMOOTH_ZG, 4, &pend.nparts, jac, cx->part_rr));
                    else
                        P4G_TRY(stencil(l, su, oc, P4G_K_SMOOTH_ZG, 4, &pend.nparts, jac));
                } else if (last && dots) {
                    P4G_TRY(pend_flush());
                    typedef OutScaleU<T, OutCheb<T, false, 2>> OC;
                    P4G_TRY(stencil(l, SrcArr<T>{b}, OC{cb, OutCheb<T, false, 2>{b, nullptr, dst, T(c0), T(c1), T(c2)}},
                                    scls, 2, &nbk, jac));
                } else {
                    if (fine) P4G_TRY(pend_flush());
                    typedef OutScaleU<T, OutCheb<T, false, 0>> OC;
                    P4G_TRY(stencil(l, SrcArr<T>{b}, OC{cb, OutCheb<T, false, 0>{b, nullptr, dst, T(c0), T(c1), T(c2)}},
                                    scls, 2, &nbk, jac));
                }
                pm1_implicit = false;
            } else if (prev) {
                if (last && dots)
                    P4G_TRY(stencil(l, SrcArr<T>{pm1}, OutCheb<T, true, 2>{b, pm2, dst, T(c0), T(c1), T(c2)},
                                    scls, 4, &nbk, jac));
                else
                    P4G_TRY(stencil(l, SrcArr<T>{pm1}, OutCheb<T, true, 0>{b, pm2, dst, T(c0), T(c1), T(c2)},
                                    scls, 4, &nbk, jac));
            } else {
                if (last && dots)
                    P4G_TRY(stencil(l, SrcArr<T>{pm1}, OutCheb<T, false, 2>{b, nullptr, dst, T(c0), T(c1), T(c2)},
                                    scls, 3, &nbk, jac));
                else
                    P4G_TRY(stencil(l, SrcArr<T>{pm1}, OutCheb<T, false, 0>{b, nullptr, dst, T(c0), T(c1), T(c2)},
                                    scls, 3, &nbk, jac));
            }
            if (nb) *nb = nbk;
            pm2 = pm1;
            pm1 = dst;
            (void)used;
        }
        return 0;
    }

    // run a pending residual update as a kernel of its own
    int pend_flush() {
        if (!pend.on) return 0;
        const long n = (*lv)[nlev - 1].n;
        const int gb = vblocks(n);
        hipStream_t s = cx->stream;
        pend.on = false;
        pend.nparts = gb;
        return cx->launch(P4G_K_CG_RESID, true, 3.0 * n * sizeof(T), [&] {
            if (pend.rr) k_r_update<T, 1><<<gb, 256, 0, s>>>(n, pend.alpha, 0, pend.r, pend.w, pend.rn, cx->part_rr);
            else k_r_update<T, 0><<<gb, 256, 0, s>>>(n, pend.alpha, 0, pend.r, pend.w, pend.rn, cx->part_rr);
        });
    }
    // whether the fine-level cycle starts with a kernel reading the residual
    // through a source policy (then the pending update is fused into it)
    bool pend_fusable(const LevelInfo& L) const {
        return fuse_r && pctype == PCT_MG && nlev > 1 && zg_fusable(L) &&
               (nu == 1 || !(nu >= 3 && smoother == SM_CHEB));
    }
    // p_new = z + beta p_old ; w = A p_new ; (p_new, w) ; x += alpha p_old
    // (init: x = alpha p_old, x not read)
    template <typename X>
    int direction(const T* zz, const T* pold, T* pnew, X* xx, const double* beta, const double* alpha, long* nbs,
                  bool init = false) {
        const long n = (*lv)[nlev - 1].n;
        const double xb = (init ? 1.0 : 2.0) * sizeof(X);   // x written (and read unless init)
        if (fuse_dir)
            return stencil(nlev - 1, SrcDirX<T, X>{zz, pold, pnew, beta, xx, alpha, init}, OutApplyDot<T>{w},
                           P4G_K_STENCIL_DOT, 4.0 + xb / sizeof(T), nbs, false);
        hipStream_t s = cx->stream;
        P4G_TRY(cx->launch(P4G_K_CG_UPDATE, true, 3.0 * n * sizeof(T) + xb * n, [&] {
            k_dir_x<T, X><<<vblocks(n), 256, 0, s>>>(n, beta, alpha, zz, pold, pnew, xx, init ? 1 : 0);
        }));
        return stencil(nlev - 1, SrcArr<T>{pnew}, OutApplyDot<T>{w}, P4G_K_STENCIL_DOT, 2, nbs, false);
    }

    // PCMGMCycle: b, x of level l (x zero on entry when xzero).  On the
    // finest level with dots, the last kernel fuses (x,b) and (x,x).
    int mcycle(int l, const T* b, T* x, bool xzero, bool dots, long* nb) {
        if (l == 0) {
            P4G_TRY(coarse(b, x, xzero));
            if (dots) {   // one-level multigrid: (x,b), (x,x) in a separate pass
                const long n = (*lv)[0].n;
                hipStream_t s = cx->stream;
                if (nb) *nb = vblocks(n);
                P4G_TRY(cx->launch(P4G_K_OTHER, true, 2.0 * n * sizeof(T), [&] {
                    k_dot2<T><<<vblocks(n), 256, 0, s>>>(n, x, b, cx->part);
                }));
            }
            return 0;
        }
        const LevelInfo& L = (*lv)[l];
        Dev& d = dv[l];
        const bool jac = jacobi != 0;
        const double idg = jac ? 1.0 / L.d : 1.0;
        const double s = smoother == SM_CHEB ? 2.0 / (L.emax + L.emin) : rich_scale;
        // pre-smoothing and residual
        bool zgf = false, restricted = false;
        T* xpre = nullptr;
        if (l == nlev - 1 && pend.on && !(xzero && pend_fusable(L))) P4G_TRY(pend_flush());
        if (xzero && nu == 1 && zg_fusable(L)) {
            zgf = true;
            if (l == nlev - 1 && pend.on) {   // b = r - alpha w, formed here and stored to b
                if (pend.rn != b) return set_error(PETSC_ERR_LIB, "internal: pending update not into b");
                const SrcUpd<T> su{pend.r, pend.w, pend.alpha, pend.rn};
                typedef OutScaleU<T, OutResid<T, 0>> OR;
                const OR orr{T(s * idg), OutResid<T, 0>{b, d.r}};
                pend.on = false;
                if (pend.rr)
                    P4G_TRY(stencil(l, su, OutRR<T, OR>{orr}, P4G_K_RESID_ZG, 4, &pend.nparts, jac, cx->part_rr));
                else
                    P4G_TRY(stencil(l, su, orr, P4G_K_RESID_ZG, 4, &pend.nparts, jac));
            } else {
                typedef OutScaleU<T, OutResid<T, 0>> OR;
                P4G_TRY(stencil(l, SrcArr<T>{b}, OR{T(s * idg), OutResid<T, 0>{b, d.r}}, P4G_K_RESID_ZG, 2, nullptr,
                                jac));
            }
        } else {
            xpre = d.t[3] ? d.t[3] : d.t[0];
            if (xpre == x && !xzero) return set_error(PETSC_ERR_LIB, "internal buffer aliasing");
            P4G_TRY(smooth(l, b, xzero ? nullptr : x, xpre, false, nullptr));
            if (rr_fusable(L)) {   // residual and restriction in one pass, the fine residual not stored
                P4G_TRY(resid_restrict(l, xpre, b, dv[l - 1].b));
                restricted = true;
            } else {
                P4G_TRY(stencil(l, SrcArr<T>{xpre}, OutResid<T, 0>{b, d.r}, P4G_K_RESID_ZG, 3, nullptr, jac));
            }
        }
        // coarse-grid correction
        Dev& c = dv[l - 1];
        if (!restricted) P4G_TRY(restrict_(l, d.r, c.b));
        for (int q = 0; q < cycles; ++q) P4G_TRY(mcycle(l - 1, c.b, c.x, q == 0, false, nullptr));
        // x = pre + P x_c, then post-smoothing into x
        T* x0 = zgf ? d.t[0] : xpre;
        const T* pre = zgf ? b : xpre;
        if (L.dim >= 2 && fuse_post && nu == 1 && !L.gal && !sor) {
            // the post-smoothing step reads the prolonged iterate on the fly
            if (dots) P4G_TRY(post_first<true>(l, zgf, pre, T(s * idg), c.x, b, x, T(s), nb));
            else P4G_TRY(post_first<false>(l, zgf, pre, T(s * idg), c.x, b, x, T(s), nb));
            return 0;
        }
        if (post2_fusable(L) && !zgf) {
            // the first of several Chebyshev post-smoothing steps fused with the
            // prolongation; t = pre + P x_c stored for the next step's recurrence
            T *tt = nullptr, *y1 = nullptr;
            for (int q = 0; q < 5; ++q) {
                T* u = d.t[q];
                if (!u || u == pre || u == b || u == x) continue;
                if (!tt) tt = u;
                else if (!y1) y1 = u;
            }
            if (!tt || !y1) return set_error(PETSC_ERR_LIB, "internal: smoother buffers exhausted");
            typedef OutKeepU<T, OutCheb<T, false, 0>> OK;
            const OK ok{tt, OutCheb<T, false, 0>{b, nullptr, y1, T(0), T(1), T(s)}};
            P4G_TRY(prolong_smooth<false>(l, pre, T(0), c.x, ok, 4.0, nullptr, jac));
            return smooth(l, b, y1, x, dots, nb, 2, tt);
        }
        P4G_TRY(prolong(l, zgf, pre, T(s * idg), c.x, x0));
        P4G_TRY(smooth(l, b, x0, x, dots, nb));
        return 0;
    }

    // prolongation + first post-smoothing step in one pass for two or more
    // Chebyshev steps (the prolonged iterate is stored for the recurrence):
    // 3D levels of the LDS-staged passes, -p4g_post_fuse auto|1|0
    bool post2_fusable(const LevelInfo& L) const {
        if (L.dim != 3 || L.var || L.gal || sor || vidiag) return false;
        if (smoother != SM_CHEB || nu < 2 || post2_mode == 0 || !lds_level(L)) return false;
        return true;
    }

    // PCMG type full (PETSc's PCMGFCycle_Private): b restricted through every
    // level, the coarsest solved from zero, then upward x_l = P x_{l-1}
    // (overwritten) and one cycle on level l from that iterate
    int fcycle(const T* b, T* x, bool dots, long* nb) {
        P4G_TRY(pend_flush());
        const int L = nlev - 1;
        P4G_TRY(restrict_(L, b, dv[L - 1].b));
        for (int l = L - 1; l > 0; --l) P4G_TRY(restrict_(l, dv[l].b, dv[l - 1].b));
        P4G_TRY(mcycle(0, dv[0].b, dv[0].x, true, false, nullptr));
        for (int l = 1; l <= L; ++l) {
            const T* bl = l == L ? b : dv[l].b;
            T* xl = l == L ? x : dv[l].x;
            P4G_TRY(prolong(l, true, bl, T(0), dv[l - 1].x, xl));   // 0 b + P x_{l-1}
            P4G_TRY(mcycle(l, bl, xl, false, l == L && dots, l == L ? nb : nullptr));
        }
        return 0;
    }

    // PCMG type kaskade (PETSc's PCMGKCycle_Private): b restricted through
    // every level, the coarsest solved from zero, then upward x_l = P x_{l-1}
    // (overwritten) and one smoother solve on level l from that iterate
    int kcycle(const T* b, T* x, bool dots, long* nb) {
        P4G_TRY(pend_flush());
        const int L = nlev - 1;
        P4G_TRY(restrict_(L, b, dv[L - 1].b));
        for (int l = L - 1; l > 0; --l) P4G_TRY(restrict_(l, dv[l].b, dv[l - 1].b));
        P4G_TRY(coarse(dv[0].b, dv[0].x));
        for (int l = 1; l <= L; ++l) {
            const T* bl = l == L ? b : dv[l].b;
            T* xl = l == L ? x : dv[l].x;
            T* t = dv[l].t[0];
            P4G_TRY(prolong(l, true, bl, T(0), dv[l - 1].x, t));   // 0 b + P x_{l-1}
            P4G_TRY(smooth(l, bl, t, xl, l == L && dots, l == L ? nb : nullptr));
        }
        return 0;
    }

    // one cycle on the finest level that stops after the pre-smoothing when
    // |b - A x| <= ttol (PCApplyRichardson_MG); *conv: the reason, or 0
    int mcycle_check(const T* b, T* x, bool xzero, double ttol, double atol, double* Sblk, int* conv) {
        const int l = nlev - 1;
        const LevelInfo& L = (*lv)[l];
        Dev& d = dv[l];
        Dev& c = dv[l - 1];
        const bool jac = jacobi != 0;
        hipStream_t st = cx->stream;
        *conv = 0;
        T* xpre = d.t[3];
        if (!xpre) return set_error(PETSC_ERR_LIB, "internal: no smoother temporaries");
        P4G_TRY(smooth(l, b, xzero ? nullptr : x, xpre, false, nullptr));
        long nbr = 0;
        P4G_TRY(stencil(l, SrcArr<T>{xpre}, OutResid<T, 1>{b, d.r}, P4G_K_RESID_ZG, 3, &nbr, jac));
        P4G_TRY(cx->finalize(nbr, Sblk, FIN_RAW, 0));
        P4G_TRY(cx->readS(cx->hS));
        const double rn = std::sqrt(cx->hS[(Sblk - cx->S) + S_TMP0]);
        if (ttol > 0.0 && rn <= ttol) {
            HIP_TRY(hipMemcpyAsync(x, xpre, L.n * sizeof(T), hipMemcpyDeviceToDevice, st));
            *conv = rn < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
            return 0;
        }
        P4G_TRY(restrict_(l, d.r, c.b));
        for (int q = 0; q < cycles; ++q) P4G_TRY(mcycle(l - 1, c.b, c.x, q == 0, false, nullptr));
        P4G_TRY(prolong(l, false, xpre, T(0), c.x, xpre));
        return smooth(l, b, xpre, x, false, nullptr);
    }

    // z = B r with fused (z,r), (z,z) partials; returns the block count
    int pc_apply_dots(const T* rr, T* zz, long* nb) {
        const LevelInfo& F = (*lv)[nlev - 1];
        if (pctype == PCT_MG && nlev > 1 && fmg) return fcycle(rr, zz, true, nb);
        if (pctype == PCT_MG && nlev > 1 && kaskade) return kcycle(rr, zz, true, nb);
        if (pctype == PCT_MG && nlev > 1 && mcycles > 1) {   // each cycle from the last iterate
            P4G_TRY(pend_flush());
            for (int c = 0; c < mcycles; ++c)
                P4G_TRY(mcycle(nlev - 1, rr, zz, c == 0, c == mcycles - 1, c == mcycles - 1 ? nb : nullptr));
            return 0;
        }
        if (pctype == PCT_MG && nlev > 1) {
            P4G_TRY(mcycle(nlev - 1, rr, zz, true, true, nb));
            if (pend.on) return set_error(PETSC_ERR_LIB, "internal: pending residual update not consumed");
            return 0;
        }
        P4G_TRY(pend_flush());
        // one level: PCMGSetLevels makes it a smoother level, applied once
        if (pctype == PCT_MG) return smooth(0, rr, nullptr, zz, true, nb);
        if (pctype == PCT_ASM) {
            if (!asm_op) return set_error(PETSC_ERR_SUP, "-pc_type asm is implemented for p4g_vi_solve");
            P4G_TRY(asm_op(rr, zz));
            const long n = F.n;
            hipStream_t s = cx->stream;
            *nb = vblocks(n);
            return cx->launch(P4G_K_OTHER, true, 2.0 * n * sizeof(T), [&] {
                k_dot2<T><<<vblocks(n), 256, 0, s>>>(n, zz, rr, cx->part);
            });
        }
        if (pctype == PCT_ILU) {
            const long n = F.n;
            hipStream_t s = cx->stream;
            const double* uf = ilu_u.size() > (size_t)(nlev - 1) ? ilu_u[nlev - 1] : nullptr;
            if (F.gal && ilu9) P4G_TRY(ilu9_run<T>(F, ilu9, rr, zz, s));
            else if (F.gal) P4G_TRY(ilu27_run<T>(F, g27.size() > (size_t)(nlev - 1) ? g27[nlev - 1] : nullptr, uf, rr, zz, s));
            else P4G_TRY(ilu_run<T>(F, uf, rr, zz, s));
            *nb = vblocks(n);
            return cx->launch(P4G_K_OTHER, true, 2.0 * n * sizeof(T), [&] {
                k_dot2<T><<<vblocks(n), 256, 0, s>>>(n, zz, rr, cx->part);
            });
        }
        const double c = (pctype == PCT_JACOBI) ? 1.0 / F.d : 1.0;
        if (pctype == PCT_JACOBI && (F.var || F.d != F.db))
            return set_error(PETSC_ERR_SUP, "-pc_type jacobi on the device needs a constant diagonal");
        *nb = vblocks(F.n);
        return scale(F.n, T(c), rr, zz, true, true);
    }

    int pc_apply(const T* rr, T* zz) {
        long nb = 0;
        P4G_TRY(pc_apply_dots(rr, zz, &nb));
        return 0;
    }

    // every per-call work array of a preconditioner application (each
    // level's b, x, r and smoother temporaries) with its length in elements;
    // f may re-point them (operator data such as kc, g27, Ainv is not here)
    template <class F>
    void each_work(F&& f) {
        for (int l = 0; l < nlev; ++l) {
            Dev& d = dv[l];
            const long n = (*lv)[l].n;
            f(d.x, n);
            f(d.b, n);
            f(d.r, n);
            for (int q = 0; q < 5; ++q) f(d.t[q], n);
        }
    }

    int mat_mult(const T* xx, T* yy) {
        return stencil(nlev - 1, SrcArr<T>{xx}, OutApply<T>{yy}, P4G_K_OTHER, 2, nullptr, false);
    }

    // KSPSolve_CG from x = 0.  history: norms of the convergence test.
    struct CGOut { int its = 0, reason = 0; double rnorm0 = 0, rnorm = 0; std::vector<double> hist; };
    int cg(const T* b, T* x, int normtype, double rtol, double atol, double dtol, int maxit,
           CGOut& o, double* Sblk) {
        const LevelInfo& F = (*lv)[nlev - 1];
        const long n = F.n;
        hipStream_t st = cx->stream;
        o = CGOut();
        // with deferred updates x is first written by the first fold (or
        // zeroed on an exit before any)
        const int D = std::is_same<T, double>::value ? x_defer : 0;
        bool xset = D == 0;
        if (xset) P4G_TRY(fill(n, T(0), x));
        if (D == 0) P4G_TRY(fill(n, T(0), pa));   // p_{-1} = 0 (with D, z stands in: beta_0 = 0)
        // r_0 = b - A 0 = b: the first iteration reads b itself, its update
        // goes to r2, and the residuals then alternate between r and r2
        const int fmode = normtype == NT_PRECOND ? FIN_BETA_PRECOND
                        : normtype == NT_NATURAL ? FIN_BETA_NATURAL : FIN_BETA_UNPRECOND;
        const int soff = (int)(Sblk - cx->S);
        if (normtype == NT_UNPRECOND) {   // (r, r) of the initial residual r = b, twice (TMP0, TMP1)
            P4G_TRY(cx->launch(P4G_K_OTHER, true, 1.0 * n * sizeof(T), [&] {
                k_dot2<T><<<vblocks(n), 256, 0, st>>>(n, b, b, cx->part);
            }));
            P4G_TRY(cx->finalize(vblocks(n), Sblk, FIN_RAW, 0));
        }
        long nb = 0;
        P4G_TRY(pc_apply_dots(b, z, &nb));
        if (normtype == NT_UNPRECOND) {
            // S_RR <- TMP1 of the raw pass above (kept by a copy on device)
            HIP_TRY(hipMemcpyAsync(Sblk + S_RR, Sblk + S_TMP1, sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        P4G_TRY(cx->finalize(nb, Sblk, fmode, 1));
        P4G_TRY(cx->readS(cx->hS));
        const double* hs = cx->hS + soff;
        double dp = hs[S_DP];
        o.rnorm0 = dp;
        o.rnorm = dp;
        o.hist.push_back(dp);
        const double ttol = std::max(rtol * dp, atol);
        if (std::isnan(dp) || std::isinf(dp)) o.reason = P4G_DIVERGED_NANORINF;
        else if (dp <= ttol) o.reason = dp < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
        else if (hs[S_RZ] == 0.0) { o.its = 1; o.reason = P4G_CONVERGED_ATOL; }
        if (o.reason) return xset ? 0 : fill(n, T(0), x);   // x = 0
        // the first direction p_0 = z + beta_0 p_{-1}, beta_0 = 0: p_{-1} is a
        // zeroed buffer, or z itself where the ring picks the direction
        // buffers (z is then never a direction pass's output)
        T* pold = D > 0 ? z : pa;
        T* pnew = pb;
        double rz_old = hs[S_RZ];   // beta of the previous iteration (its sign is tested)
        const T* rc = b;   // current residual; the update r - alpha w goes to rx
        T* rx = r2;
        double dpi_old = 0.0;
        // with x_defer (FP64) the direction pass leaves x alone: p_i stays in
        // a ring of x_defer + 2 buffers (a listed buffer is not reused) with
        // alpha_i copied to S[32 + j], and the listed updates are applied in
        // iteration order when the list is full and on exit (the same FP64
        // rounding as applying each in the next direction pass)
        FoldListT<T> fl{};
        double* Sal = cx->S + 32;
        std::vector<T*> ring = {pa, pb};
        for (int j = 0; j < D; ++j) ring.push_back(pdef[j]);
        auto listed = [&](const T* p) {
            for (int j = 0; j < fl.m; ++j)
                if (fl.p[j] == p) return true;
            return false;
        };
        auto fold = [&]() -> int {
            if constexpr (std::is_same<T, double>::value) {
                P4G_TRY(launch_fold_x(cx, n, x, fl, !xset));
                xset = true;
                return 0;
            }
            return set_error(PETSC_ERR_LIB, "internal: deferred updates need an FP64 iterate");
        };
        // the current direction and its alpha listed (the list folded first when full)
        auto list_current = [&]() -> int {
            if (fl.m == D) P4G_TRY(fold());
            P4G_TRY(cx->launch(P4G_K_OTHER, false, 0.0, [&] {
                k_copy_scalar<<<1, 64, 0, st>>>(Sblk + S_ALPHA, Sal + fl.m);
            }));
            fl.p[fl.m] = pold;
            fl.a[fl.m] = Sal + fl.m;
            ++fl.m;
            return 0;
        };
        // the next direction (x += alpha_i p_i applied or listed)
        auto next_direction = [&](bool first) -> int {
            long nbd = 0;
            if (!D) {
                P4G_TRY(direction<T>(z, pold, pnew, x, Sblk + S_BCOEF, Sblk + S_ALPHA, &nbd));
                return cx->finalize(nbd, Sblk, FIN_ALPHA, 0);
            }
            if (!first) P4G_TRY(list_current());
            pnew = nullptr;
            for (T* q : ring)
                if (q != pold && !listed(q)) { pnew = q; break; }
            if (!pnew) return set_error(PETSC_ERR_LIB, "internal: direction buffers exhausted");
            P4G_TRY(stencil(nlev - 1, SrcDir<T>{z, pold, pnew, Sblk + S_BCOEF}, OutApplyDot<T>{w}, P4G_K_STENCIL_DOT,
                            4, &nbd, false));
            return cx->finalize(nbd, Sblk, FIN_ALPHA, 0);
        };
        // x += alpha p is deferred into the next direction update; alpha = 0
        // before the first iteration.  The direction update of iteration i+1
        // needs no host decision (beta is formed on the device), so it is
        // queued before the host waits for iteration i's scalars: the
        // convergence test overlaps it, and on exit it has applied the last
        // deferred update, so x is complete.
        k_fill<double><<<1, 256, 0, st>>>(1, 0.0, Sblk + S_ALPHA);
        LAUNCH_CHECK();
        P4G_TRY(next_direction(true));
        int i = 0;
        for (; i < maxit; ++i) {
            o.its = i + 1;
            std::swap(pold, pnew);   // pold now holds the current direction
            // r - alpha w, fused into the first kernel of the preconditioner
            pend.r = rc;
            pend.w = w;
            pend.alpha = Sblk + S_ALPHA;
            pend.rn = rx;
            pend.rr = normtype == NT_UNPRECOND;
            pend.on = true;
            P4G_TRY(pc_apply_dots(rx, z, &nb));
            T* const next = rc == b ? r : const_cast<T*>(rc);   // b is never written
            rc = rx;
            rx = next;
            if (normtype == NT_UNPRECOND) P4G_TRY(cx->finalize(pend.nparts, Sblk, FIN_RR, 0, cx->part_rr));
            P4G_TRY(cx->finalize(nb, Sblk, fmode, 0));
            // the next direction is queued before the host reads this
            // iteration's scalars, unless (deferred updates) the norm is
            // predicted, from the last two, to meet the tolerance now: then
            // the host tests first, so no pass is spent on a direction that
            // a converged solve does not use
            const size_t nh = o.hist.size();
            const double pred = nh >= 2 ? o.hist[nh - 1] * (o.hist[nh - 1] / o.hist[nh - 2]) : o.hist[nh - 1];
            const bool spec = D == 0 || pred > ttol;
            P4G_TRY(cx->readS_begin(cx->hS));
            if (spec) P4G_TRY(next_direction(false));   // next direction; x += alpha_i p_i applied or listed
            P4G_TRY(cx->readS_end());
            const double dpi = hs[S_PW];
            const bool indefinite = dpi == 0.0 || (i > 0 && ((dpi > 0) != (dpi_old > 0)));
            if (!spec && !indefinite) {
                const double dn = hs[S_DP];
                if (std::isnan(dn) || std::isinf(dn) || dn <= ttol) {
                    P4G_TRY(list_current());   // this iteration's update, no next direction
                    o.rnorm = dn;
                    o.hist.push_back(dn);
                    o.reason = std::isnan(dn) || std::isinf(dn) ? P4G_DIVERGED_NANORINF
                             : dn < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
                    break;
                }
                P4G_TRY(next_direction(false));
                P4G_TRY(cx->readS(cx->hS));
            }
            if (indefinite) {
                // PETSc stops before x += alpha_i p_i: a listed update is
                // dropped (it is the last one listed); with x_defer 0 the
                // direction pass has already applied it
                if (spec && D > 0) --fl.m;
                o.reason = P4G_DIVERGED_INDEFINITE_MAT;
                break;
            }
            dpi_old = dpi;
            dp = hs[S_DP];
            o.rnorm = dp;
            o.hist.push_back(dp);
            if (std::isnan(dp) || std::isinf(dp)) { o.reason = P4G_DIVERGED_NANORINF; break; }
            if (dp <= ttol) { o.reason = dp < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL; break; }
            if (dp >= dtol * o.rnorm0) { o.reason = P4G_DIVERGED_DTOL; break; }
            // beta = 0, or beta changing sign: PETSc stops at the top of the next iteration
            const double rz = hs[S_RZ];
            if (rz == 0.0 && i + 1 < maxit) {
                o.its = i + 2;
                o.reason = P4G_CONVERGED_ATOL;
                break;
            }
            if (rz * rz_old < 0.0 && i + 1 < maxit) {
                o.its = i + 2;
                o.reason = P4G_DIVERGED_INDEFINITE_PC;
                break;
            }
            rz_old = rz;
        }
        if (i >= maxit && !o.reason) o.reason = P4G_DIVERGED_ITS;
        if (fl.m) P4G_TRY(fold());   // the listed updates: x complete on every exit
        if (!xset) P4G_TRY(fill(n, T(0), x));
        return 0;
    }

    // KSPSolve_GMRES from x = 0: restarted GMRES(m) with left preconditioning,
    // classical Gram-Schmidt without refinement and the preconditioned
    // residual norm of the Givens-rotated Hessenberg least-squares problem.
    // A cycle starts from the true residual B (b - A x); the convergence test
    // runs at each cycle start and after each Arnoldi step; hist records the
    // norms where PETSc's KSPGMRESCycle calls its monitor.  The basis, its
    // dots and the updates stay on the device; the host forms the rotations
    // from the it + 2 scalars of each step.
    T* gm_V = nullptr;
    double *gm_c = nullptr, *gm_part = nullptr;
    int gm_m = 0;
    int gmres(const T* b, T* x, double rtol, double atol, double dtol, int maxit, int m, CGOut& o, double* Sblk) {
        const LevelInfo& F = (*lv)[nlev - 1];
        const long n = F.n;
        hipStream_t st = cx->stream;
        const int nb = vblocks(n);
        o = CGOut();
        if (m < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-ksp_gmres_restart must be >= 1");
        if (!gm_V || gm_m < m) {
            P4G_TRY(alloc(&gm_V, (long)(m + 1) * n));
            double* c = nullptr;
            HIP_TRY(hipMalloc((void**)&c, (size_t)(m + 1) * sizeof(double)));
            allocs.push_back(c);
            gm_c = c;
            HIP_TRY(hipMalloc((void**)&c, (size_t)(m + 1) * nb * sizeof(double)));
            allocs.push_back(c);
            gm_part = c;
            gm_m = m;
        }
        T* V = gm_V;
        const int soff = (int)(Sblk - cx->S);
        const double* hs = cx->hS + soff;
        const int ld = m + 2;
        std::vector<double> H((size_t)ld * (m + 1), 0.0), cc(m + 1), ss(m + 1), grs(m + 2), hh(m + 2);
        auto HH = [&](int a, int c) -> double& { return H[(size_t)c * ld + a]; };
        // |v|^2 into Sblk[S_TMP0] and v /= |v| (VecNormalize)
        auto normalize = [&](T* v) -> int {
            k_dot2<T><<<nb, 256, 0, st>>>(n, v, v, cx->part);
            LAUNCH_CHECK();
            P4G_TRY(cx->finalize(nb, Sblk, FIN_RAW, 0));
            k_normalize<<<nb, 256, 0, st>>>(n, Sblk + S_TMP0, v, v);
            LAUNCH_CHECK();
            return 0;
        };
        P4G_TRY(fill(n, T(0), x));
        double ttol = 0.0, res = 0.0;
        int its = 0, reason = 0;
        bool first = true;
        while (!reason) {
            if (first) {
                P4G_TRY(pc_apply(b, V));
            } else {
                P4G_TRY(stencil(nlev - 1, SrcArr<T>{x}, OutResid<T, 0>{b, r}, P4G_K_OTHER, 3, nullptr, false));
                P4G_TRY(pc_apply(r, V));
            }
            P4G_TRY(normalize(V));
            P4G_TRY(cx->readS(cx->hS));
            res = std::sqrt(hs[S_TMP0]);
            grs[0] = res;
            o.hist.push_back(res);
            if (first) { o.rnorm0 = res; ttol = std::max(rtol * res, atol); }
            first = false;
            if (res == 0.0) { reason = P4G_CONVERGED_ATOL; break; }
            auto test = [&](int k, double rn) -> int {
                if (std::isnan(rn) || std::isinf(rn)) return P4G_DIVERGED_NANORINF;
                if (rn <= ttol) return rn < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
                if (k > 0 && rn >= dtol * o.rnorm0) return P4G_DIVERGED_DTOL;
                return 0;
            };
            reason = test(its, res);
            int it = 0;
            bool hapend = false;
            while (!reason && it < m && its < maxit) {
                if (it) o.hist.push_back(res);
                T* vi = V + (long)it * n;
                T* vn = V + (long)(it + 1) * n;
                P4G_TRY(mat_mult(vi, w));
                P4G_TRY(pc_apply(w, vn));
                k_mdot<<<nb, 256, 0, st>>>(n, it + 1, V, vn, gm_part);
                k_mfinal<<<it + 1, 256, 0, st>>>(nb, gm_part, gm_c);
                k_maxpy<<<nb, 256, 0, st>>>(n, it + 1, -1.0, gm_c, V, vn);
                LAUNCH_CHECK();
                HIP_TRY(hipMemcpyAsync(hh.data(), gm_c, (it + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
                P4G_TRY(normalize(vn));
                P4G_TRY(cx->readS(cx->hS));
                const double tt = std::sqrt(hs[S_TMP0]);
                for (int j = 0; j <= it; ++j) HH(j, it) = hh[j];
                HH(it + 1, it) = tt;
                double hapbnd = std::fabs(tt / grs[it]);
                if (hapbnd > 1e-30) hapbnd = 1e-30;
                if (tt < hapbnd) hapend = true;
                for (int j = 1; j <= it; ++j) {   // the earlier rotations
                    const double a0 = HH(j - 1, it), a1 = HH(j, it);
                    HH(j - 1, it) = cc[j - 1] * a0 + ss[j - 1] * a1;
                    HH(j, it) = cc[j - 1] * a1 - ss[j - 1] * a0;
                }
                if (!hapend) {
                    const double a0 = HH(it, it), a1 = HH(it + 1, it);
                    const double rr = std::sqrt(a0 * a0 + a1 * a1);
                    if (rr == 0.0) { reason = P4G_DIVERGED_NULL; break; }
                    cc[it] = a0 / rr;
                    ss[it] = a1 / rr;
                    grs[it + 1] = -(ss[it] * grs[it]);
                    grs[it] = cc[it] * grs[it];
                    HH(it, it) = cc[it] * a0 + ss[it] * a1;
                    res = std::fabs(grs[it + 1]);
                } else {
                    res = 0.0;
                }
                ++it;
                ++its;
                reason = test(its, res);
                if (hapend && !reason) { reason = P4G_DIVERGED_BREAKDOWN; break; }
            }
            if (it && (reason || its >= maxit)) o.hist.push_back(res);
            if (it > 0) {   // KSPGMRESBuildSoln: back substitution, x += V y
                const int top = it - 1;
                std::vector<double> y(grs.begin(), grs.begin() + it);
                if (HH(top, top) != 0.0) y[top] = grs[top] / HH(top, top);
                else { reason = P4G_DIVERGED_BREAKDOWN; y[top] = 0.0; }
                for (int ii = 1; ii <= top; ++ii) {
                    const int kk = top - ii;
                    double sum = grs[kk];
                    for (int j = kk + 1; j <= top; ++j) sum -= HH(kk, j) * y[j];
                    if (HH(kk, kk) == 0.0) { reason = P4G_DIVERGED_BREAKDOWN; break; }
                    y[kk] = sum / HH(kk, kk);
                }
                HIP_TRY(hipMemcpyAsync(gm_c, y.data(), it * sizeof(double), hipMemcpyHostToDevice, st));
                k_maxpy<<<nb, 256, 0, st>>>(n, it, 1.0, gm_c, V, x);
                LAUNCH_CHECK();
                HIP_TRY(hipStreamSynchronize(st));   // y lives on the host stack
            }
            if (its >= maxit && !reason) reason = P4G_DIVERGED_ITS;
        }
        o.its = its;
        o.reason = reason;
        o.rnorm = res;
        return 0;
    }

    // KSPSolve_Richardson from x = 0.  monitor: PETSc's generic loop (taken
    // when a monitor is set, or the PC is not multigrid): the norm of r
    // (unpreconditioned) or B r (preconditioned) is monitored and tested,
    // then x += scale B r, r = b - A x; after maxit updates one more norm and
    // test.  Otherwise, with multigrid, PCApplyRichardson_MG: up to maxit
    // cycles, each stopping after the finest pre-smoothing once
    // |b - A x| <= max(rtol |b|, atol); its = the completed cycles.
    int richardson(const T* b, T* x, int normtype, double rtol, double atol, double dtol, int maxit,
                   double scale, bool monitor, CGOut& o, double* Sblk) {
        const LevelInfo& F = (*lv)[nlev - 1];
        const long n = F.n;
        hipStream_t st = cx->stream;
        const int gb = vblocks(n);
        const double* hs = cx->hS + (Sblk - cx->S);
        o = CGOut();
        P4G_TRY(fill(n, T(0), x));
        auto norm2 = [&](const T* v, double& nv) -> int {
            P4G_TRY(cx->launch(P4G_K_OTHER, true, double(n) * sizeof(T), [&] {
                k_dot2<T><<<gb, 256, 0, st>>>(n, v, v, cx->part);
            }));
            P4G_TRY(cx->finalize(gb, Sblk, FIN_RAW, 0));
            P4G_TRY(cx->readS(cx->hS));
            nv = std::sqrt(hs[S_TMP0]);
            return 0;
        };
        if (pctype == PCT_MG && nlev > 1 && !monitor) {
            double ttol = 0.0;
            if (rtol > 0.0) {
                double bn;
                P4G_TRY(norm2(b, bn));
                ttol = std::max(rtol * bn, atol);
            } else if (atol > 0.0) {
                ttol = atol;
            }
            int i = 0, conv = 0;
            for (; i < maxit && !conv; ++i) {
                P4G_TRY(mcycle_check(b, x, i == 0, ttol, atol, Sblk, &conv));
                if (conv) break;
            }
            o.its = i;
            o.reason = conv ? conv : P4G_CONVERGED_ITS;
            return 0;
        }
        HIP_TRY(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        auto pcz = [&](double& zn) -> int {   // z = B r and |z|
            long nb = 0;
            P4G_TRY(pc_apply_dots(r, z, &nb));
            P4G_TRY(cx->finalize(nb, Sblk, FIN_RAW, 0));
            P4G_TRY(cx->readS(cx->hS));
            zn = std::sqrt(hs[S_TMP1]);
            return 0;
        };
        double ttol = 0.0, rn = 0.0;
        auto test = [&](int it) -> int {   // KSPConvergedDefault
            if (it == 0) { o.rnorm0 = rn; ttol = std::max(rtol * rn, atol); }
            if (std::isnan(rn) || std::isinf(rn)) return P4G_DIVERGED_NANORINF;
            if (rn <= ttol) return rn < atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
            if (it > 0 && rn >= dtol * o.rnorm0) return P4G_DIVERGED_DTOL;
            return 0;
        };
        const bool unp = normtype == NT_UNPRECOND;
        int i = 0, its = 0;
        for (; i < maxit; ++i) {
            if (unp) P4G_TRY(norm2(r, rn));
            else P4G_TRY(pcz(rn));
            o.rnorm = rn;
            o.hist.push_back(rn);
            o.reason = test(i);
            if (o.reason) break;
            double zn;
            if (unp) P4G_TRY(pcz(zn));
            P4G_TRY(cx->launch(P4G_K_CG_UPDATE, true, 3.0 * n * sizeof(T), [&] {
                k_axpy<T><<<gb, 256, 0, st>>>(n, T(scale), z, x);
            }));
            ++its;
            P4G_TRY(stencil(nlev - 1, SrcArr<T>{x}, OutResid<T, 0>{b, r}, P4G_K_CG_RESID, 3, nullptr, false));
        }
        if (!o.reason) {
            if (unp) P4G_TRY(norm2(r, rn));
            else P4G_TRY(pcz(rn));
            o.rnorm = rn;
            o.hist.push_back(rn);
            if (its >= maxit) {
                o.reason = test(i);
                if (!o.reason) o.reason = P4G_DIVERGED_ITS;
            }
        }
        o.its = its;
        return 0;
    }
};

// FP64 Lanczos (PCG) estimate of the extreme eigenvalues of D^-1 A on one
// level: the same procedure as the oracle's KSPChebyshev estimate.
static void tridiag_extreme(int m, const double* a, const double* b, double* emin, double* emax) {
    double T[16][16];
    std::memset(T, 0, sizeof T);
    for (int i = 0; i < m; i++) {
        T[i][i] = a[i];
        if (i + 1 < m) T[i][i + 1] = T[i + 1][i] = b[i];
    }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) off += T[p][q] * T[p][q];
        if (off < 1e-30) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                if (std::fabs(T[p][q]) < 1e-300) continue;
                const double th = (T[q][q] - T[p][p]) / (2.0 * T[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < m; r++) {
                    const double trp = T[r][p], trq = T[r][q];
                    T[r][p] = c * trp - s * trq;
                    T[r][q] = s * trp + c * trq;
                }
                for (int r = 0; r < m; r++) {
                    const double tpr = T[p][r], tqr = T[q][r];
                    T[p][r] = c * tpr - s * tqr;
                    T[q][r] = s * tpr + c * tqr;
                }
            }
    }
    *emin = *emax = T[0][0];
    for (int i = 1; i < m; i++) {
        *emin = std::min(*emin, T[i][i]);
        *emax = std::max(*emax, T[i][i]);
    }
}

// r -= a w ; z = id r (id array or constant) ; acc0 += r.z
__global__ void __launch_bounds__(256)
k_lanczos_update(long n, double a, double* __restrict__ r, const double* __restrict__ w,
                 const double* __restrict__ idv, double idc, double* __restrict__ z,
                 double* __restrict__ part) {
    double acc = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double rv = r[q] - a * w[q];
        r[q] = rv;
        const double zv = rv * (idv ? idv[q] : idc);
        z[q] = zv;
        acc += rv * zv;
    }
    __shared__ double sm[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
        part[2 * blockIdx.x + 1] = 0.0;
    }
}

struct OutIdiag {   // stores the Jacobi factor of every row
    double* __restrict__ id;
    static constexpr int NDOT = 0, NPRE = 0;
    __device__ __forceinline__ void point(long q, double, double, double idg, double&, double&) const {
        id[q] = idg;
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<double, V>&, const TV<double, V>&,
                                           const TV<double, V>& idg, double&, double&) const {
        stv<double, V>(id + q, idg);
    }
};

}  // namespace p4g

// ---------------------------------------------------------------- solver

struct p4g_solver_s {
    p4g_da_s* da = nullptr;
    Options opt;
    std::string opts_str;
    // configuration
    int pctype = PCT_MG, nlev = 1, cycles = 1, smoother = SM_CHEB, nu = 2, jacobi = 1;
    double rich_scale = 1.0;
    int esteig_steps = 10;
    int normtype = NT_PRECOND, maxit = 10000;
    double rtol = 1e-5, atol = 1e-50, dtol = 1e5;
    bool mixed = false;
    double reliable_delta = 1e-1;
    bool fp32_diff = true;    // FP32 engine: difference-form rows
    bool mixed_x64 = true;    // mixed: accumulate the iterate in FP64 every iteration
    int x_defer = -1;         // CG iterate updates deferred (at most; -1: by path)
    float* pdef[kFoldMax] = {};   //   their direction buffers
    int fuse_r = -1, fuse_dir = -1;   // CG updates fused into stencil passes: 1 on, 0 off, -1 by rule
    bool galerkin = false;            // coarse operators P^T A P (-pc_mg_galerkin)
    bool fmg = false;                 // -pc_mg_type full: one F-cycle per preconditioner apply
    bool kaskade = false;             // -pc_mg_type kaskade
    int mcycles = 1;                  // -pc_mg_multiplicative_cycles
    bool ksp_rich = false;            // -ksp_type richardson
    bool ksp_gmres = false;           // -ksp_type gmres
    int gmres_restart = 30;           // -ksp_gmres_restart
    double rich_outer_scale = 1.0;    // -ksp_richardson_scale
    bool monitor = false;             // -ksp_monitor (PETSc's generic Richardson loop)
    bool sor = false;                 // -mg_levels_pc_type sor|ilu|icc: explicit level smoothers
    bool lilu = false;                //   ... with ILU(0) instead of SSOR
    std::vector<double*> ilu_u;       // per level: ILU(0) pivots (device) or null
    int coarse_its = 0, coarse_pc = 0;   // -mg_coarse_ksp_type cg: iterations and PC (0 none, 1 jacobi, 2 ilu)
    int sor_blocks = 1;               // -p4g_sor_blocks: local sweeps of that many ranks
    int lds = -1;                     // -p4g_lds auto|1|0: LDS-staged 3D stencil passes
    int resid_restrict = -1;          // -p4g_resid_restrict auto|1|0: fused residual + restriction
    int rr_rows = 8;                  // -p4g_rr_rows 8|4: its coarse rows per block
    int rr_planes = 0;                // -p4g_rr_planes n: its coarse planes per block (0: by size)
    int asm_overlap = 1;              // -pc_asm_overlap (PCASM, VI solves)
    int ranks = 1;                    // -p4g_ranks: the processes whose DMDA split PCASM's subdomains follow
    // PCASM's subdomains on the device: per owned row its subdomain's index
    // list and inverse row (rebuilt every Newton step of a VI solve)
    int* asm_idx = nullptr;           // subdomain vertex indices, concatenated
    double* asm_inv = nullptr;        // subdomain inverses, row-major, concatenated
    long* asm_row = nullptr;          // per owned row: [inverse row offset, index offset, count, vertex]
    long asm_nrows = 0;
    size_t asm_cap_idx = 0, asm_cap_inv = 0, asm_cap_row = 0;
    int post_fuse = -1;               // -p4g_post_fuse auto|1|0: prolongation + first of several post steps
    bool sor_test_stall = false;      // -p4g_sor_test_stall: every pipelined-SOR wait gives up at once
    bool sor_pipeline = true;         // -p4g_sor_pipeline 1|0: 3D SOR by pipelined strips (else hyperplanes)
    // reduced-space VI Newton (p4g_vi_solve): SNES options and state
    double snes_rtol = 1e-8, snes_atol = 1e-50, snes_stol = 1e-8, snes_divtol = 1e4;
    int snes_maxit = 50;
    std::string snes_type, ls_type;
    bool vi_mode = false;                 // set up for p4g_vi_solve (or p4g_ms_solve: operators per Newton step)
    bool ms_mode = false;                 // set up for p4g_ms_solve
    double* ms_buf = nullptr;             // its vectors
    double* ms_ilu9 = nullptr;            // its box-pattern ILU(0) factor (-pc_type ilu)
    std::vector<SorPipe> ms_pipes;        // its levels' pipelined SOR flags (-mg_levels_pc_type sor)
    double* ms_band = nullptr;            // the coarse band LU factor, uploaded for the inverse
    long ms_band_cap = 0;
    std::vector<double> ms_fhist;         // |F| per Newton iterate
    std::vector<int> ms_kits, ms_kreason, ms_sym;   // per Newton step
    double* est_scratch = nullptr;        // device: eigen-estimate vectors, est_cap doubles (VI solves)
    long est_cap = 0;
    double* vi_buf = nullptr;             // device: F0 X F Y W G Fm xl xu, n doubles each
    unsigned char* vi_act = nullptr;      // device: active flags
    double* vi_g27 = nullptr;             // device: the finest null-row operator (27 planes)
    std::vector<double> vi_fhist;         // |F| on the inactive set, iterations 0..its
    std::vector<int> vi_kits, vi_kreason; // per Newton step: KSP iterations and reason
    // state
    bool ready = false;
    std::vector<LevelInfo> lv;
    std::vector<double> Ainv;   // dense inverse of the coarsest operator (host)
    std::vector<double> hist;   // norms of the convergence test of the last solve, from iteration 0
    std::vector<double*> g27;   // per level: 27-point Galerkin coefficients (device, FP64) or null
    int n0 = 0;
    Ctx cx;
    std::unique_ptr<Engine<double>> e64;
    std::unique_ptr<Engine<float>> e32;
    // mixed-precision outer vectors; FP64 fine cells of a variable operator
    double *xa = nullptr, *xb = nullptr, *kc64 = nullptr;
    float *r32 = nullptr, *e32v = nullptr;
    std::vector<void*> allocs;
    p4g_Report rep;

    ~p4g_solver_s() {
        e64.reset();
        e32.reset();
        for (void* p : allocs) hipFree(p);
        if (cx.part) hipFree(cx.part);
        if (cx.part_rr) hipFree(cx.part_rr);
        if (cx.S) hipFree(cx.S);
        if (cx.hS) hipHostFree(cx.hS);
        if (cx.stream) hipStreamDestroy(cx.stream);
    }
};

// deferred FP64 iterate updates of the mixed-precision PCG (-p4g_x_defer, default 6)
static inline int mixed_defer(const p4g_solver_s* s) { return s->x_defer < 0 ? 6 : s->x_defer; }

namespace p4g {

static int dalloc(p4g_solver_s* s, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
    if (e != hipSuccess) return set_error(PETSC_ERR_MEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    s->allocs.push_back(*p);
    return 0;
}

static bool is_bnd(const LevelInfo& L, int i, int j, int k) {
    bool b = i == 0 || i == L.mx - 1;
    if (L.dim > 1) b = b || j == 0 || j == L.my - 1;
    if (L.dim > 2) b = b || k == 0 || k == L.mz - 1;
    return b;
}

// host row of the level operator (FP64), for the coarse factorization
static void host_row(const LevelInfo& L, int i, int j, int k, double& dg, double s[6]) {
    for (int q = 0; q < 6; ++q) s[q] = 0.0;
    if (!L.var) {
        if (is_bnd(L, i, j, k)) { dg = L.db; return; }
        dg = L.d;
        s[0] = (i - 1 > 0) ? L.sx : 0; s[1] = (i + 1 < L.mx - 1) ? L.sx : 0;
        if (L.dim > 1) { s[2] = (j - 1 > 0) ? L.sy : 0; s[3] = (j + 1 < L.my - 1) ? L.sy : 0; }
        if (L.dim > 2) { s[4] = (k - 1 > 0) ? L.sz : 0; s[5] = (k + 1 < L.mz - 1) ? L.sz : 0; }
        return;
    }
    auto cell = [&](int a, int b, int c) { return L.kcell[((long)c * L.cy + b) * L.cx + a]; };
    if (is_bnd(L, i, j, k)) {
        double sum = 0.0;
        int cnt = 0;
        for (int dk = -1; dk <= 0; ++dk)
            for (int dj = -1; dj <= 0; ++dj)
                for (int di = -1; di <= 0; ++di) {
                    int a = i + di, b = j + dj, c = k + dk;
                    if (a < 0 || b < 0 || c < 0 || a >= L.cx || b >= L.cy || c >= L.cz) continue;
                    sum += cell(a, b, c);
                    ++cnt;
                }
        dg = 2.0 * (L.ax + L.ay + L.az) * (sum / cnt);
        return;
    }
    const double c000 = cell(i - 1, j - 1, k - 1), c100 = cell(i, j - 1, k - 1);
    const double c010 = cell(i - 1, j, k - 1), c110 = cell(i, j, k - 1);
    const double c001 = cell(i - 1, j - 1, k), c101 = cell(i, j - 1, k);
    const double c011 = cell(i - 1, j, k), c111 = cell(i, j, k);
    const double kw = 0.25 * (c000 + c010 + c001 + c011) * L.ax;
    const double ke = 0.25 * (c100 + c110 + c101 + c111) * L.ax;
    const double ks = 0.25 * (c000 + c100 + c001 + c101) * L.ay;
    const double kn = 0.25 * (c010 + c110 + c011 + c111) * L.ay;
    const double kd = 0.25 * (c000 + c100 + c010 + c110) * L.az;
    const double ku = 0.25 * (c001 + c101 + c011 + c111) * L.az;
    dg = kw + ke + ks + kn + kd + ku;
    s[0] = (i - 1 > 0) ? kw : 0; s[1] = (i + 1 < L.mx - 1) ? ke : 0;
    s[2] = (j - 1 > 0) ? ks : 0; s[3] = (j + 1 < L.my - 1) ? kn : 0;
    s[4] = (k - 1 > 0) ? kd : 0; s[5] = (k + 1 < L.mz - 1) ? ku : 0;
}

// dense Cholesky of the coarsest operator and the inverse of its factor;
// g27: the operator's 27-point coefficients (device), else its 5/7-point rows
static int coarse_factor(const LevelInfo& L, const double* g27, std::vector<double>& Ainv) {
    const int n = (int)L.n;
    if (n > MAX_COARSE)
        return set_error(PETSC_ERR_SUP, "coarsest grid has %d points (> %d); use more multigrid levels",
                         n, MAX_COARSE);
    std::vector<double> A((size_t)n * n, 0.0);
    const int mx = L.mx, my = L.my;
    if (g27) {
        std::vector<double> g((size_t)27 * n);
        HIP_TRY(hipMemcpy(g.data(), g27, g.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < L.mz; ++k)
            for (int j = 0; j < L.my; ++j)
                for (int i = 0; i < L.mx; ++i) {
                    const long q = ((long)k * my + j) * mx + i;
                    for (int c = 0; c < 27; ++c) {
                        const int oz = c / 9 - 1, oy = (c / 3) % 3 - 1, ox = c % 3 - 1;
                        if (i + ox < 0 || i + ox >= mx || j + oy < 0 || j + oy >= my || k + oz < 0 || k + oz >= L.mz)
                            continue;
                        A[q * n + q + ((long)oz * my + oy) * mx + ox] = g[(size_t)c * n + q];
                    }
                }
        // a null row (an active vertex of a reduced-space VI step) has a zero
        // row and column: factor it as an identity row, so x_q = b_q
        for (int q = 0; q < n; ++q)
            if (A[(size_t)q * n + q] == 0.0) A[(size_t)q * n + q] = 1.0;
    }
    for (int k = 0; k < (g27 ? 0 : L.mz); ++k)
        for (int j = 0; j < L.my; ++j)
            for (int i = 0; i < L.mx; ++i) {
                const long q = ((long)k * my + j) * mx + i;
                double dg, s[6];
                host_row(L, i, j, k, dg, s);
                A[q * n + q] = dg;
                if (s[0] != 0) A[q * n + q - 1] = -s[0];
                if (s[1] != 0) A[q * n + q + 1] = -s[1];
                if (s[2] != 0) A[q * n + q - mx] = -s[2];
                if (s[3] != 0) A[q * n + q + mx] = -s[3];
                if (s[4] != 0) A[q * n + q - (long)mx * my] = -s[4];
                if (s[5] != 0) A[q * n + q + (long)mx * my] = -s[5];
            }
    // A = L L^T (lower, in place)
    for (int c = 0; c < n; ++c) {
        double dsum = A[(size_t)c * n + c];
        for (int k = 0; k < c; ++k) dsum -= A[(size_t)c * n + k] * A[(size_t)c * n + k];
        if (!(dsum > 0.0)) return set_error(PETSC_ERR_LIB, "coarse Cholesky: matrix not positive definite");
        const double lcc = std::sqrt(dsum);
        A[(size_t)c * n + c] = lcc;
        for (int r = c + 1; r < n; ++r) {
            double v = A[(size_t)r * n + c];
            for (int k = 0; k < c; ++k) v -= A[(size_t)r * n + k] * A[(size_t)c * n + k];
            A[(size_t)r * n + c] = v / lcc;
        }
    }
    // Linv = L^-1 by forward substitution on the identity
    std::vector<double> Linv((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        for (int r = c; r < n; ++r) {
            double v = (r == c) ? 1.0 : 0.0;
            for (int k = c; k < r; ++k) v -= A[(size_t)r * n + k] * Linv[(size_t)k * n + c];
            Linv[(size_t)r * n + c] = v / A[(size_t)r * n + r];
        }
    }
    // A^-1 = L^-T L^-1 (symmetric): entry (r,c) sums k >= max(r,c)
    Ainv.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {               // rows of L^-1 are contiguous
        const double* Lk = &Linv[(size_t)k * n];
        for (int r = 0; r <= k; ++r) {
            const double a = Lk[r];
            if (a == 0.0) continue;
            double* Ar = &Ainv[(size_t)r * n];
            for (int c = 0; c <= r; ++c) Ar[c] += a * Lk[c];
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < r; ++c) Ainv[(size_t)c * n + r] = Ainv[(size_t)r * n + c];
    return 0;
}

// -pc_mg_galerkin: the coarse operators P^T A P, formed on the device in
// 27-point form level by level from the finest (rediscretized) operator
static int galerkin_levels(p4g_solver_s* s) {
    if ((int)s->g27.size() != s->nlev) s->g27.assign(s->nlev, nullptr);
    if (!s->galerkin) return 0;
    if (s->lv.back().var)
        return set_error(PETSC_ERR_SUP, "-pc_mg_galerkin with a variable coefficient is not implemented on the device");
    hipStream_t st = s->cx.stream;
    for (int l = s->nlev - 2; l >= 0; --l) {
        LevelInfo& C = s->lv[l];
        const LevelInfo& L = s->lv[l + 1];
        if (!s->g27[l]) P4G_TRY(dalloc(s, (void**)&s->g27[l], 27 * C.n * sizeof(double)));
        GalFine f;
        f.dim = L.dim; f.mx = L.mx; f.my = L.my; f.mz = L.mz;
        f.d = L.d; f.db = L.db; f.sx = L.sx; f.sy = L.sy; f.sz = L.sz;
        f.g = s->g27[l + 1];   // null on a rediscretized finest level
        f.n = L.n;
        k_galerkin<<<vblocks(C.n), 256, 0, st>>>(f, C.mx, C.my, C.mz, s->g27[l]);
        LAUNCH_CHECK();
        C.gal = true;
    }
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

static int estimate_level(p4g_solver_s* s, int l) {
    LevelInfo& L = s->lv[l];
    const long n = L.n;
    hipStream_t st = s->cx.stream;
    std::vector<void*> tmp;
    long used = 0;   // slices of the solver's estimate scratch when it has one (VI solves)
    auto get = [&](double** p) -> int {
        if (s->est_scratch && used + n <= s->est_cap) {
            *p = s->est_scratch + used;
            used += n;
            return 0;
        }
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, n * sizeof(double));
        if (e != hipSuccess) return set_error(PETSC_ERR_MEM, "hipMalloc (estimate): %s", hipGetErrorString(e));
        tmp.push_back(q);
        *p = (double*)q;
        return 0;
    };
    struct Free { std::vector<void*>* t; ~Free() { for (void* q : *t) hipFree(q); } } fr{&tmp};
    double *r, *z, *pa, *pb, *w, *idv = nullptr, *kc = nullptr;
    P4G_TRY(get(&r)); P4G_TRY(get(&z)); P4G_TRY(get(&pa)); P4G_TRY(get(&pb)); P4G_TRY(get(&w));
    const bool jac = s->jacobi != 0;
    OpConst<double> oc;
    OpVar<double> ov;
    const dim3 g = stencil_grid<double>(L.dim, L.mx, L.my, L.mz), bk = stencil_block(L.dim);
    const long nbs = L.gal ? vblocks(n) : nblocks(g);
    const Op27<double> og{L.gal ? s->g27[l] : nullptr, n, L.mx, L.my, L.mz, jac ? 1 : 0};
    if (L.gal) {
        P4G_TRY(get(&idv));
        k_stencil27<double, SrcZero<double>, OutIdiag><<<vblocks(n), 256, 0, st>>>(og, SrcZero<double>{}, OutIdiag{idv},
                                                                                 s->cx.part);
        LAUNCH_CHECK();
    } else if (!L.var) {
        oc.d = L.d; oc.db = L.db; oc.sx = L.sx; oc.sy = L.sy; oc.sz = L.sz;
        oc.id = jac ? 1.0 / L.d : 1.0; oc.idb = jac ? 1.0 / L.db : 1.0;
    } else {
        const long nc = (long)L.cx * L.cy * L.cz;
        void* q = nullptr;
        HIP_TRY(hipMalloc(&q, nc * sizeof(double)));
        tmp.push_back(q);
        kc = (double*)q;
        HIP_TRY(hipMemcpyAsync(kc, L.kcell.data(), nc * sizeof(double), hipMemcpyHostToDevice, st));
        ov.kc = kc; ov.cx = L.cx; ov.cy = L.cy; ov.cz = L.cz;
        ov.ax = L.ax; ov.ay = L.ay; ov.az = L.az; ov.jac = jac ? 1 : 0;
        P4G_TRY(get(&idv));
        k_stencil<double, 3, OpVar<double>, SrcZero<double>, OutIdiag>
            <<<g, bk, 0, st>>>(L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), ov, SrcZero<double>{}, OutIdiag{idv}, s->cx.part);
        LAUNCH_CHECK();
    }
    const double idc = jac ? 1.0 / L.d : 1.0;
    const int gb = vblocks(n);
    double* S = s->cx.S + 16;
    // r = noisy, z = D^-1 r, rz = (r,z), p = 0 (so p_new = z at beta = 0)
    k_noisy<<<gb, 256, 0, st>>>(n, r);
    k_fill<double><<<gb, 256, 0, st>>>(n, 0.0, pa);
    HIP_TRY(hipMemsetAsync(S, 0, 16 * sizeof(double), st));
    k_fill<double><<<gb, 256, 0, st>>>(n, 0.0, w);
    // r -= alpha w, z = B r (B = D^-1, or one SSOR sweep), partial (r, z)
    auto lupdate = [&](double alpha) -> int {
        if (s->sor) {
            if (alpha != 0.0) k_axpy<double><<<gb, 256, 0, st>>>(n, -alpha, w, r);
            if (s->lilu) P4G_TRY(ilu_run<double>(L, s->ilu_u.size() > (size_t)l ? s->ilu_u[l] : nullptr, r, z, st));
            else P4G_TRY(sor_run<double>(L, s->g27.size() > (size_t)l ? s->g27[l] : nullptr, r, z, st));
            k_dot2<double><<<gb, 256, 0, st>>>(n, z, r, s->cx.part);
        } else {
            k_lanczos_update<<<gb, 256, 0, st>>>(n, alpha, r, w, idv, idc, z, s->cx.part);
        }
        LAUNCH_CHECK();
        return 0;
    };
    P4G_TRY(lupdate(0.0));
    k_finalize<<<1, FT, 0, st>>>(gb, s->cx.part, S, FIN_RAW, 0);
    P4G_TRY(s->cx.readS(s->cx.hS));
    double rz = s->cx.hS[16 + S_TMP0], rzold = 0.0;
    double al[16], be[16];
    int m = 0;
    double* pold = pa;
    double* pnew = pb;
    for (int it = 0; it < s->esteig_steps; ++it) {
        if (!(rz > 0.0)) break;
        double beta = 0.0;
        if (it > 0) { beta = rz / rzold; be[it - 1] = beta; }
        s->cx.hS[16 + S_BCOEF] = beta;
        HIP_TRY(hipMemcpyAsync(S + S_BCOEF, &s->cx.hS[16 + S_BCOEF], sizeof(double), hipMemcpyHostToDevice, st));
        if (s->e64 && s->e64->mf_op && l == s->nlev - 1) {   // matrix-free: p = z + beta p, w = A p, (p, w)
            k_comb3<double><<<gb, 256, 0, st>>>(n, 0.0, nullptr, beta, pold, 1.0, z, pnew);
            LAUNCH_CHECK();
            P4G_TRY(s->e64->mf_op(pnew, w));
            k_dot2<double><<<gb, 256, 0, st>>>(n, pnew, w, s->cx.part);
        } else if (L.gal)
            k_stencil27<double, SrcDir<double>, OutApplyDot<double>><<<vblocks(n), 256, 0, st>>>(
                og, SrcDir<double>{z, pold, pnew, S + S_BCOEF}, OutApplyDot<double>{w}, s->cx.part);
        else if (L.var)
            k_stencil<double, 3, OpVar<double>, SrcDir<double>, OutApplyDot<double>><<<g, bk, 0, st>>>(
                L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), ov, SrcDir<double>{z, pold, pnew, S + S_BCOEF}, OutApplyDot<double>{w}, s->cx.part);
        else if (L.dim == 3)
            k_stencil<double, 3, OpConst<double>, SrcDir<double>, OutApplyDot<double>><<<g, bk, 0, st>>>(
                L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), oc, SrcDir<double>{z, pold, pnew, S + S_BCOEF}, OutApplyDot<double>{w}, s->cx.part);
        else if (L.dim == 2)
            k_stencil<double, 2, OpConst<double>, SrcDir<double>, OutApplyDot<double>><<<g, bk, 0, st>>>(
                L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), oc, SrcDir<double>{z, pold, pnew, S + S_BCOEF}, OutApplyDot<double>{w}, s->cx.part);
        else
            k_stencil<double, 1, OpConst<double>, SrcDir<double>, OutApplyDot<double>><<<g, bk, 0, st>>>(
                L.mx, L.my, L.mz, kchf(L.dim, L.my, L.mz), oc, SrcDir<double>{z, pold, pnew, S + S_BCOEF}, OutApplyDot<double>{w}, s->cx.part);
        LAUNCH_CHECK();
        std::swap(pold, pnew);
        k_finalize<<<1, FT, 0, st>>>(nbs, s->cx.part, S, FIN_RAW, 0);
        P4G_TRY(s->cx.readS(s->cx.hS));
        const double pw = s->cx.hS[16 + S_TMP0];
        if (!(pw > 0.0)) break;
        const double alpha = rz / pw;
        al[it] = alpha;
        m = it + 1;
        P4G_TRY(lupdate(alpha));
        k_finalize<<<1, FT, 0, st>>>(gb, s->cx.part, S, FIN_RAW, 0);
        P4G_TRY(s->cx.readS(s->cx.hS));
        rzold = rz;
        rz = s->cx.hS[16 + S_TMP0];
        if (rz <= 1e-28 * std::fabs(rzold)) break;
    }
    double a[16], b[16], emin = 1.0, emax = 1.0;
    for (int j = 0; j < m; ++j) {
        a[j] = 1.0 / al[j] + (j > 0 ? be[j - 1] / al[j - 1] : 0.0);
        if (j + 1 < m) b[j] = std::sqrt(be[j]) / al[j];
    }
    if (m > 0) tridiag_extreme(m, a, b, &emin, &emax);
    L.emin = 0.0 * emin + 0.1 * emax;
    L.emax = 0.0 * emin + 1.1 * emax;
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

static const char* kUnsupported[] = {"mg_levels_ksp_norm_type", nullptr};

static int configure(p4g_solver_s* s) {
    Options& o = s->opt;
    const std::string kt = o.gets("ksp_type", "cg");
    if (kt == "richardson") s->ksp_rich = true;
    else if (kt == "gmres") s->ksp_gmres = true;
    else if (kt != "cg")
        return set_error(PETSC_ERR_SUP, "-ksp_type %s: the device path implements cg, gmres and richardson",
                         kt.c_str());
    s->gmres_restart = o.geti("ksp_gmres_restart", 30);
    if (s->gmres_restart < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-ksp_gmres_restart must be >= 1");
    s->rich_outer_scale = o.getr("ksp_richardson_scale", 1.0);
    s->monitor = o.has("ksp_monitor");
    const std::string pt = o.gets("pc_type", "");
    if (pt == "mg") s->pctype = PCT_MG;
    else if (pt == "jacobi") s->pctype = PCT_JACOBI;
    else if (pt == "none") s->pctype = PCT_NONE;
    else if (pt.empty() || pt == "ilu" || pt == "icc") s->pctype = PCT_ILU;   // PETSc's serial default
    else if (pt == "asm") s->pctype = PCT_ASM;   // VI solves (p4g_vi_solve) only
    else return set_error(PETSC_ERR_SUP, "-pc_type %s is not implemented on the device", pt.c_str());
    if (s->pctype == PCT_ASM) {
        if (o.gets("sub_pc_type", "ilu") != "lu")
            return set_error(PETSC_ERR_SUP, "-pc_type asm: the subdomain solves are exact (-sub_pc_type lu)");
        s->asm_overlap = o.geti("pc_asm_overlap", 1);
        if (s->asm_overlap < 0) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-pc_asm_overlap must be >= 0");
    }
    s->ranks = o.geti("p4g_ranks", 1);
    if (s->ranks < 1 || s->ranks > 64) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_ranks must be 1..64");
    for (const char** u = kUnsupported; *u; ++u)
        if (o.has(*u)) return set_error(PETSC_ERR_SUP, "-%s is not implemented on the device", *u);
    const std::string mt = o.gets("pc_mg_type", "multiplicative");
    if (mt == "full") s->fmg = true;
    else if (mt == "kaskade") s->kaskade = true;
    else if (mt != "multiplicative")
        return set_error(PETSC_ERR_SUP,
                         "-pc_mg_type %s is not implemented on the device (multiplicative, full, kaskade)", mt.c_str());
    s->mcycles = o.geti("pc_mg_multiplicative_cycles", 1);
    if (s->mcycles < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-pc_mg_multiplicative_cycles must be >= 1");
    s->snes_rtol = o.getr("snes_rtol", 1e-8);
    s->snes_atol = o.getr("snes_atol", 1e-50);
    s->snes_stol = o.getr("snes_stol", 1e-8);
    s->snes_divtol = o.getr("snes_divergence_tolerance", 1e4);
    s->snes_maxit = o.geti("snes_max_it", 50);
    s->snes_type = o.gets("snes_type", "");
    s->ls_type = o.gets("snes_linesearch_type", "bt");
    s->rtol = o.getr("ksp_rtol", 1e-5);
    s->atol = o.getr("ksp_atol", 1e-50);
    s->dtol = o.getr("ksp_divtol", 1e5);
    s->maxit = o.geti("ksp_max_it", 10000);
    const std::string nt = o.gets("ksp_norm_type", "preconditioned");
    if (nt == "preconditioned") s->normtype = NT_PRECOND;
    else if (nt == "unpreconditioned") s->normtype = NT_UNPRECOND;
    else if (nt == "natural") s->normtype = NT_NATURAL;
    else return set_error(PETSC_ERR_SUP, "-ksp_norm_type %s not supported", nt.c_str());
    s->nlev = s->pctype == PCT_MG ? o.geti("pc_mg_levels", s->da->refine + 1) : 1;
    if (s->nlev < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-pc_mg_levels must be >= 1");
    const std::string ct = o.gets("pc_mg_cycle_type", "v");
    s->cycles = ct == "w" ? 2 : 1;
    if (ct != "v" && ct != "w") return set_error(PETSC_ERR_ARG_WRONG, "-pc_mg_cycle_type %s", ct.c_str());
    const std::string lk = o.gets("mg_levels_ksp_type", "chebyshev");
    if (lk == "chebyshev") s->smoother = SM_CHEB;
    else if (lk == "richardson") s->smoother = SM_RICH;
    else return set_error(PETSC_ERR_SUP, "-mg_levels_ksp_type %s not implemented on the device", lk.c_str());
    const std::string lp = o.gets("mg_levels_pc_type", s->pctype == PCT_MG ? "sor" : "jacobi");
    if (lp == "jacobi") s->jacobi = 1;
    else if (lp == "none") s->jacobi = 0;
    else if ((lp == "ilu" || lp == "icc") && s->pctype == PCT_MG) {   // ILU(0) = IC(0) for this operator
        s->sor = true;
        s->lilu = true;
        s->jacobi = 1;
    }
    else if (lp == "sor" && s->pctype == PCT_MG) {   // PETSc's default: symmetric sweep, omega 1
        s->sor = true;
        s->jacobi = 1;
        for (const char* k : {"mg_levels_pc_sor_forward", "mg_levels_pc_sor_backward",
                              "mg_levels_pc_sor_local_forward", "mg_levels_pc_sor_local_backward"})
            if (o.has(k)) return set_error(PETSC_ERR_SUP, "-%s: the device SOR smoother sweeps symmetrically", k);
        if (o.getr("mg_levels_pc_sor_omega", 1.0) != 1.0 || o.geti("mg_levels_pc_sor_its", 1) != 1 ||
            o.geti("mg_levels_pc_sor_lits", 1) != 1)
            return set_error(PETSC_ERR_SUP, "the device SOR smoother has omega 1 and one sweep");
        s->sor_blocks = o.geti("p4g_sor_blocks", 1);
        s->sor_test_stall = o.getb("p4g_sor_test_stall", false);
        s->sor_pipeline = o.getb("p4g_sor_pipeline", true);
        if (s->sor_blocks < 1 || s->sor_blocks > 16)
            return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_sor_blocks must be in 1..16");
    }
    else if (s->nlev == 1) s->jacobi = 1;   /* one level: no smoother is used */
    else
        return set_error(PETSC_ERR_SUP, "-mg_levels_pc_type %s is not implemented on the device (PETSc's "
                                        "default is sor); use -mg_levels_pc_type jacobi", lp.c_str());
    s->nu = o.geti("mg_levels_ksp_max_it", 2);
    if (s->nu < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-mg_levels_ksp_max_it must be >= 1");
    s->rich_scale = o.getr("mg_levels_ksp_richardson_scale", 1.0);
    s->esteig_steps = std::max(1, std::min(15, o.geti("mg_levels_ksp_chebyshev_esteig_steps", 10)));
    const std::string cks = o.gets("mg_coarse_ksp_type", "preonly");
    const std::string cpc = o.gets("mg_coarse_pc_type", "lu");
    if (cks == "cg" && s->pctype == PCT_MG) {   // c/ch6/study/coarseperformance.sh: fixed-count PCG
        if (o.gets("mg_coarse_ksp_convergence_test", "") != "skip")
            return set_error(PETSC_ERR_SUP, "-mg_coarse_ksp_type cg runs a fixed count on the device: add "
                                            "-mg_coarse_ksp_convergence_test skip");
        s->coarse_its = o.geti("mg_coarse_ksp_max_it", 10000);
        if (s->coarse_its < 1) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-mg_coarse_ksp_max_it must be >= 1");
        if (cpc == "icc" || cpc == "ilu") s->coarse_pc = 2;
        else if (cpc == "jacobi") s->coarse_pc = 1;
        else if (cpc == "none") s->coarse_pc = 0;
        else return set_error(PETSC_ERR_SUP, "-mg_coarse_pc_type %s with a CG coarse solve (icc, ilu, jacobi, none)",
                              cpc.c_str());
    } else if (cks != "preonly" || (cpc != "lu" && cpc != "cholesky" && cpc != "redundant")) {
        return set_error(PETSC_ERR_SUP, "coarse solver must be preonly + lu/cholesky/redundant or cg (fixed count)");
    }
    const std::string pr = o.gets("p4g_precision", "double");
    if (pr == "mixed") s->mixed = true;
    else if (pr != "double") return set_error(PETSC_ERR_ARG_WRONG, "-p4g_precision %s", pr.c_str());
    s->reliable_delta = o.getr("p4g_reliable_delta", 1e-1);
    s->fp32_diff = o.getb("p4g_fp32_diff_form", true);
    s->mixed_x64 = o.getb("p4g_mixed_x64", true);
    s->x_defer = o.geti("p4g_x_defer", -1);   // -1: 6 (mixed), by memory (FP64 CG of KSP solves)
    if (s->x_defer < -1 || s->x_defer > kFoldMax - 1)
        return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_x_defer must be in 0..%d", kFoldMax - 1);
    {
        auto tri = [&](const std::string& f, const char* name, int& v) -> int {
            if (f == "auto") v = -1;
            else if (f == "1" || f == "true" || f == "yes") v = 1;
            else if (f == "0" || f == "false" || f == "no") v = 0;
            else return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-%s must be auto, 1 or 0", name);
            return 0;
        };
        P4G_TRY(tri(o.gets("p4g_lds", "auto"), "p4g_lds", s->lds));
        P4G_TRY(tri(o.gets("p4g_resid_restrict", "auto"), "p4g_resid_restrict", s->resid_restrict));
        s->rr_rows = o.geti("p4g_rr_rows", 8);
        if (s->rr_rows != 4 && s->rr_rows != 8)
            return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_rr_rows must be 4 or 8");
        s->rr_planes = o.geti("p4g_rr_planes", 0);
        if (s->rr_planes < 0 || s->rr_planes > 64)
            return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_rr_planes must be in 0..64");
        P4G_TRY(tri(o.gets("p4g_post_fuse", "auto"), "p4g_post_fuse", s->post_fuse));
        const std::string both = o.gets("p4g_fuse_cg", "auto");
        P4G_TRY(tri(both, "p4g_fuse_cg", s->fuse_r));
        P4G_TRY(tri(o.gets("p4g_fuse_r", both), "p4g_fuse_r", s->fuse_r));
        P4G_TRY(tri(o.gets("p4g_fuse_dir", both), "p4g_fuse_dir", s->fuse_dir));
    }
    if (!(s->reliable_delta > 0.0 && s->reliable_delta < 1.0))
        return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-p4g_reliable_delta must be in (0,1)");
    if (s->mixed && s->pctype != PCT_MG) return set_error(PETSC_ERR_SUP, "-p4g_precision mixed needs -pc_type mg");
    if (o.has("pc_mg_galerkin")) {   // PETSc: both | pmat | mat | none; the bare flag means both
        const std::string g = o.gets("pc_mg_galerkin", "");
        if (g.empty() || g == "both" || g == "pmat" || g == "mat" || g == "true" || g == "1")
            s->galerkin = s->pctype == PCT_MG && s->nlev > 1;
        else if (g != "none" && g != "false" && g != "0")
            return set_error(PETSC_ERR_ARG_WRONG, "-pc_mg_galerkin %s", g.c_str());
    }
    for (const char* vw : {"ksp_view_mat", "ksp_view_rhs", "ksp_view_solution"})
        if (o.has(vw)) {
            const std::string v = o.gets(vw, "");
            if (v.compare(0, 7, "binary:") != 0 || v.size() == 7)
                return set_error(PETSC_ERR_SUP, "-%s %s: the device path writes binary:<file> viewers only", vw,
                                 v.c_str());
        }
    if (s->galerkin && s->mixed)
        return set_error(PETSC_ERR_SUP, "-pc_mg_galerkin with -p4g_precision mixed is not implemented on the device");
    if ((s->pctype == PCT_ILU || s->lilu) && o.geti("pc_factor_levels", 0) + o.geti("mg_levels_pc_factor_levels", 0) != 0)
        return set_error(PETSC_ERR_SUP, "the device ILU is ILU(0): -pc_factor_levels 0");
    if ((s->pctype == PCT_ILU || s->lilu) && s->galerkin)
        return set_error(PETSC_ERR_SUP, "ILU(0) of Galerkin coarse levels is not implemented on the device");
    if (s->pctype == PCT_ILU && (s->mixed || !s->da->kcell.empty()))
        return set_error(PETSC_ERR_SUP, "-pc_type ilu needs FP64 and a constant-coefficient operator on the device");
    if (s->coarse_its > 0 && (s->mixed || s->galerkin || !s->da->kcell.empty() || s->nlev < 2))
        return set_error(PETSC_ERR_SUP, "-mg_coarse_ksp_type cg needs FP64, a rediscretized constant-coefficient "
                                        "coarsest level and at least two levels on the device");
    if (s->sor && (s->mixed || !s->da->kcell.empty()))
        return set_error(PETSC_ERR_SUP, "-mg_levels_pc_type sor needs FP64 and a constant-coefficient operator "
                                        "on the device");
    if (s->ksp_gmres && (s->mixed || s->normtype != NT_PRECOND))
        return set_error(PETSC_ERR_SUP, "-ksp_type gmres runs in FP64 with the preconditioned norm (left "
                                        "preconditioning) on the device");
    if (s->ksp_rich && s->mixed)
        return set_error(PETSC_ERR_SUP, "-ksp_type richardson with -p4g_precision mixed is not implemented on the device");
    if (s->ksp_rich && s->normtype == NT_NATURAL)
        return set_error(PETSC_ERR_SUP, "-ksp_norm_type natural is not supported by -ksp_type richardson");
    if ((s->kaskade || s->mcycles > 1) && s->mixed)
        return set_error(PETSC_ERR_SUP, "-pc_mg_type kaskade and -pc_mg_multiplicative_cycles with -p4g_precision "
                                        "mixed are not implemented on the device");
    return 0;
}

// classify the rows the Jacobian callback inserts on one level
static int extract_level(p4g_solver_s* s, LevelInfo& L, p4g_da_s* lda) {
    p4g_da_s* da = s->da;
    L.dim = lda->dim; L.mx = lda->mx; L.my = lda->dim > 1 ? lda->my : 1; L.mz = lda->dim > 2 ? lda->mz : 1;
    L.n = (long)L.mx * L.my * L.mz;
    for (int d = 0; d < 3; ++d) L.h[d] = d < L.dim ? lda->h(d) : 0.0;
    if (!da->kcell.empty()) return 0;   // variable coefficient handled by the caller
    if (!da->jac) return set_error(PETSC_ERR_ARG_WRONG, "no Jacobian callback: call p4g_da_set_jacobian_local");
    p4g_mat_s M;
    M.da = lda;
    M.seen.assign(L.n, 0);
    std::vector<double> u(L.n, 0.0);
    std::vector<double*> rows;
    std::vector<double**> planes;
    void* au = make_view(lda, u.data(), rows, planes);
    DMDALocalInfo info;
    info_of(lda, &info);
    int e = da->jac(&info, au, &M, &M, da->jacctx);
    if (e) return e;
    if (M.rows != L.n)
        return set_error(PETSC_ERR_ARG_WRONG, "Jacobian callback set %ld of %ld rows on the %dx%dx%d grid",
                         M.rows, L.n, L.mx, L.my, L.mz);
    if (M.bad)
        return set_error(PETSC_ERR_SUP, "Jacobian on the %dx%dx%d grid is not a constant-coefficient star "
                         "stencil with eliminated Dirichlet rows (%ld rows differ; first: %s)",
                         L.mx, L.my, L.mz, M.bad, M.why.c_str());
    L.var = false;
    L.d = M.have_i ? M.d : M.db;
    L.db = M.have_b ? M.db : M.d;
    L.sx = M.s[0]; L.sy = M.s[1]; L.sz = M.s[2];
    return 0;
}

static int build_levels(p4g_solver_s* s) {
    p4g_da_s* da = s->da;
    const int nlev = s->nlev;
    s->lv.assign(nlev, LevelInfo());
    std::vector<p4g_da_s> grids(nlev);
    grids[nlev - 1] = *da;
    grids[nlev - 1].internal = true;
    for (int l = nlev - 1; l > 0; --l) {
        p4g_da_s c = grids[l];
        int m[3] = {c.mx, c.my, c.mz};
        for (int d = 0; d < c.dim; ++d) {
            if (m[d] < 3 || (m[d] - 1) % 2)
                return set_error(PETSC_ERR_ARG_OUTOFRANGE, "grid %dx%dx%d cannot be coarsened for %d levels",
                                 da->mx, da->my, da->mz, nlev);
            m[d] = (m[d] - 1) / 2 + 1;
        }
        c.mx = m[0]; c.my = m[1]; c.mz = m[2];
        c.kcell.clear();
        grids[l - 1] = c;
    }
    for (int l = 0; l < nlev; ++l) {
        grids[l].kcell.clear();
        P4G_TRY(extract_level(s, s->lv[l], &grids[l]));
    }
    if (!da->kcell.empty()) {
        if (da->dim != 3) return set_error(PETSC_ERR_SUP, "cell coefficients need a 3D grid");
        std::vector<double> kc = da->kcell;
        for (int l = nlev - 1; l >= 0; --l) {
            LevelInfo& L = s->lv[l];
            L.var = true;
            L.cx = L.mx - 1; L.cy = L.my - 1; L.cz = L.mz - 1;
            const double dvol = L.h[0] * L.h[1] * L.h[2];
            L.ax = dvol / (L.h[0] * L.h[0]); L.ay = dvol / (L.h[1] * L.h[1]); L.az = dvol / (L.h[2] * L.h[2]);
            L.kcell = kc;
            L.d = L.db = 1.0;   // unused by the variable operator
            if (l > 0) {   // coarse cell = mean of its 8 fine cells
                const int cx = L.cx / 2, cy = L.cy / 2, cz = L.cz / 2;
                std::vector<double> c((size_t)cx * cy * cz);
                for (int k = 0; k < cz; ++k)
                    for (int j = 0; j < cy; ++j)
                        for (int i = 0; i < cx; ++i) {
                            double sum = 0.0;
                            for (int dk = 0; dk < 2; ++dk)
                                for (int dj = 0; dj < 2; ++dj)
                                    for (int di = 0; di < 2; ++di)
                                        sum += kc[((size_t)(2 * k + dk) * L.cy + (2 * j + dj)) * L.cx + (2 * i + di)];
                            c[((size_t)k * cy + j) * cx + i] = 0.125 * sum;
                        }
                kc.swap(c);
            }
        }
    }
    return 0;
}

// Which CG updates run inside stencil passes (the residual update in the
// first kernel of the cycle, the direction update with the operator apply).
// The fused passes read two arrays at stencil neighbours; the split passes
// stream, and re-read what is still in the 256 MB last-level cache.
// Measured per solve: 257^3 FP64 7.58 ms split / 7.70 residual fused /
// 7.90 both; 513^3 FP32 (mixed) 53.6 / 54.0 / 54.1; 513^3 FP64 94.3 split
// residual / 93.7 both.  Both are fused above 768 MiB per fine vector.
template <typename T>
static void fuse_rules(const p4g_solver_s* s, Engine<T>* e, double vec_bytes) {
    // by rule: the residual update once a vector exceeds the 256 MB last-level
    // cache (513^3: FP64 and FP32), the direction update from 768 MiB
    const bool beyond_llc = vec_bytes > 256.0 * 1048576.0, big = vec_bytes > 768.0 * 1048576.0;
    e->fuse_r = s->fuse_r >= 0 ? s->fuse_r == 1 : beyond_llc;
    e->fuse_dir = s->fuse_dir >= 0 ? s->fuse_dir == 1 : big;
}

// local SOR sweeps of several ranks: the slowest dimension of the finest
// grid split as M/np + (M mod np > q); a coarse index I belongs to the rank
// owning fine index 2I (PETSc's DMDA ownership under refinement)
static void sor_splits(p4g_solver_s* s) {
    const int np = s->sor_blocks;
    for (LevelInfo& L : s->lv) { L.sor_nblk = 1; L.sor_blk[0] = 0; }
    if (np <= 1) return;
    auto slow = [](const LevelInfo& L) { return L.dim == 3 ? L.mz : (L.dim == 2 ? L.my : L.mx); };
    const int F = s->nlev - 1;
    std::vector<std::vector<int>> sp(s->nlev, std::vector<int>(np + 1, 0));
    const int M = slow(s->lv[F]);
    for (int q = 0; q < np; ++q) sp[F][q + 1] = sp[F][q] + M / np + (M % np > q ? 1 : 0);
    for (int l = F; l > 0; --l) {
        const int mc = slow(s->lv[l - 1]);
        for (int q = 0; q < np; ++q) {
            int cnt = 0;
            for (int I = 0; I < mc; ++I)
                if (2 * I >= sp[l][q] && 2 * I < sp[l][q + 1]) ++cnt;
            sp[l - 1][q + 1] = sp[l - 1][q] + cnt;
        }
    }
    for (int l = 0; l <= F; ++l) {
        s->lv[l].sor_nblk = np;
        for (int q = 0; q < np; ++q) s->lv[l].sor_blk[q] = sp[l][q];
    }
}

int solver_setup(p4g_solver_s* s) {
    if (s->ready) return 0;
    const double t0 = now_ms();
    P4G_TRY(build_levels(s));
    sor_splits(s);
    // pipelined 3D SOR sweeps (k_sor7_strips): progress flags per level
    if (s->pctype == PCT_MG && s->sor && !s->lilu && s->sor_pipeline && !s->vi_mode && !s->ms_mode) {
        s->ms_pipes.resize(s->nlev);
        for (int l = 1; l < s->nlev; ++l) {
            LevelInfo& L = s->lv[l];
            if (L.dim != 3 || L.sor_nblk > 1) continue;
            const long nf = (long)((L.my + 63) / 64) * L.mz;
            SorPipe& P = s->ms_pipes[l];
            const size_t bytes = (nf + 1) * sizeof(unsigned long long) + sizeof(int);
            P4G_TRY(dalloc(s, (void**)&P.flags, bytes));
            HIP_TRY(hipMemset(P.flags, 0, bytes));
            P.err = (int*)(P.flags + nf + 1);
            L.pipe = &P;
        }
    }
    // ILU(0) pivots: the finest level for -pc_type ilu, the smoothing levels
    // for -mg_levels_pc_type ilu (computed once on the host, in the factor's
    // own row order)
    s->ilu_u.assign(s->nlev, nullptr);
    for (int l = 0; l < s->nlev; ++l) {
        const bool need = (s->pctype == PCT_ILU && l == s->nlev - 1) || (s->lilu && (l > 0 || s->nlev == 1)) ||
                          (l == 0 && s->coarse_its > 0 && s->coarse_pc == 2);
        if (!need) continue;
        const LevelInfo& L = s->lv[l];
        if (L.var || L.gal) return set_error(PETSC_ERR_SUP, "ILU(0) is implemented for constant 3/5/7-point levels");
        std::vector<double> u;
        ilu_pivots(L, u);
        P4G_TRY(dalloc(s, (void**)&s->ilu_u[l], L.n * sizeof(double)));
        HIP_TRY(hipMemcpy(s->ilu_u[l], u.data(), L.n * sizeof(double), hipMemcpyHostToDevice));
    }
    if (!s->cx.stream) HIP_TRY(hipStreamCreate(&s->cx.stream));
    // partial-sum capacity: the largest grid of blocks of any fused pass
    const LevelInfo& F = s->lv.back();
    long cap = std::max<long>(VBLOCKS, max_part_blocks(s->lv));
    if (!s->cx.part) {
        HIP_TRY(hipMalloc(&s->cx.part, 2 * cap * sizeof(double)));
        HIP_TRY(hipMalloc(&s->cx.part_rr, 2 * cap * sizeof(double)));
        s->cx.npart_cap = cap;
        HIP_TRY(hipMalloc(&s->cx.S, 64 * sizeof(double)));   // 32..47: saved step lengths
        double init[64] = {0};
        init[S_ONE] = init[16 + S_ONE] = 1.0;
        HIP_TRY(hipMemcpy(s->cx.S, init, sizeof init, hipMemcpyHostToDevice));
        HIP_TRY(hipHostMalloc((void**)&s->cx.hS, 32 * sizeof(double), hipHostMallocDefault));
    }
    if (s->pctype == PCT_MG && s->vi_mode) {
        // VI solves form the operators, coarse factor and bounds per Newton
        // step from the active set; here only the sizes
        s->g27.assign(s->nlev, nullptr);
        s->n0 = s->nlev > 1 ? (int)s->lv[0].n : 1;
        s->Ainv.assign((size_t)s->n0 * s->n0, 0.0);
    } else if (s->pctype == PCT_MG) {
        P4G_TRY(galerkin_levels(s));
        if (s->nlev > 1 && s->coarse_its == 0) {
            P4G_TRY(coarse_factor(s->lv[0], s->g27[0], s->Ainv));
            s->n0 = (int)s->lv[0].n;
        } else {   // one level (a smoother level) or an iterative coarse solve: no dense factor
            s->n0 = 1;
            s->Ainv.assign(1, 0.0);
        }
        if (s->smoother == SM_CHEB)
            for (int l = s->nlev > 1 ? 1 : 0; l < s->nlev; ++l) P4G_TRY(estimate_level(s, l));
    }
    auto cfg = [&](auto& e) {
        e->pctype = s->pctype; e->cycles = s->cycles; e->smoother = s->smoother; e->nu = s->nu;
        e->jacobi = s->jacobi; e->rich_scale = s->rich_scale;
        e->g27.assign(s->g27.begin(), s->g27.end());
        e->fmg = s->fmg;
        e->vi = s->vi_mode;
        e->kaskade = s->kaskade;
        e->mcycles = s->mcycles;
        e->full_temps = s->kaskade || s->mcycles > 1 || (s->ksp_rich && s->pctype == PCT_MG);
        e->sor = s->sor;
        e->lilu = s->lilu;
        e->coarse_its = s->coarse_its;
        e->coarse_pc = s->coarse_pc;
        e->ilu_u.assign(s->ilu_u.begin(), s->ilu_u.end());
        e->lds = s->lds;
        e->resid_restrict_mode = s->resid_restrict;
        e->rr_rows = s->rr_rows;
        e->rr_planes = s->rr_planes;
        e->post2_mode = s->post_fuse;
    };
    if (F.var) {
        const size_t nc = F.kcell.size();
        P4G_TRY(dalloc(s, (void**)&s->kc64, nc * sizeof(double)));
        HIP_TRY(hipMemcpy(s->kc64, F.kcell.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    }
    if (s->mixed) {
        s->e32.reset(new Engine<float>);
        cfg(s->e32);
        s->e32->diff_form = s->fp32_diff;
        fuse_rules(s, s->e32.get(), F.n * sizeof(float));
        P4G_TRY(s->e32->setup(&s->cx, &s->lv, s->Ainv, s->n0));
        const long n = F.n;
        P4G_TRY(dalloc(s, (void**)&s->xa, n * sizeof(double)));
        P4G_TRY(dalloc(s, (void**)&s->xb, n * sizeof(double)));
        P4G_TRY(dalloc(s, (void**)&s->r32, n * sizeof(float)));
        P4G_TRY(dalloc(s, (void**)&s->e32v, n * sizeof(float)));
        if (s->mixed_x64)
            for (int j = 0; j < mixed_defer(s); ++j) P4G_TRY(dalloc(s, (void**)&s->pdef[j], n * sizeof(float)));
    } else {
        s->e64.reset(new Engine<double>);
        cfg(s->e64);
        fuse_rules(s, s->e64.get(), F.n * sizeof(double));
        if (s->vi_mode) s->e64->fuse_r = s->e64->fuse_dir = false;
        // KSP solves (not the Newton solves of p4g_vi_solve / p4g_ms_solve)
        // (default: as many as 16 GiB of FP64 directions hold, at most 15;
        // none below 3, where the folds move as much as the per-iteration updates)
        const long dbytes = F.n * (long)sizeof(double);
        int dauto = (int)std::min<long>(kFoldMax - 1, (16L << 30) / std::max(1L, dbytes));
        if (dauto < 3) dauto = 0;
        s->e64->x_defer = s->vi_mode || s->ms_mode ? 0 : (s->x_defer < 0 ? dauto : s->x_defer);
        P4G_TRY(s->e64->setup(&s->cx, &s->lv, s->Ainv, s->n0));
    }
    HIP_TRY(hipStreamSynchronize(s->cx.stream));
    s->ready = true;
    s->rep.nlevels = s->nlev;
    s->rep.emax_fine = s->nlev > 1 ? s->lv.back().emax : 0.0;
    s->rep.setup_ms = now_ms() - t0;
    return 0;
}

// one FP64 stencil pass on the finest level (outer residual, MatMult)
template <class Src, class Out>
static int fine64(p4g_solver_s* s, Src src, Out out, int cls, double bytes, long* nb) {
    const LevelInfo& F = s->lv.back();
    const dim3 g = stencil_grid<double>(F.dim, F.mx, F.my, F.mz), bk = stencil_block(F.dim);
    if (nb) *nb = nblocks(g);
    hipStream_t st = s->cx.stream;
    double* part = s->cx.part;
    if (F.var) {
        OpVar<double> ov;
        ov.kc = s->kc64; ov.cx = F.cx; ov.cy = F.cy; ov.cz = F.cz;
        ov.ax = F.ax; ov.ay = F.ay; ov.az = F.az; ov.jac = 0;
        return s->cx.launch(cls, true, bytes, [&] {
            k_stencil<double, 3, OpVar<double>, Src, Out><<<g, bk, 0, st>>>(F.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz), ov, src, out, part);
        });
    }
    OpConst<double> oc;
    oc.d = F.d; oc.db = F.db; oc.sx = F.sx; oc.sy = F.sy; oc.sz = F.sz; oc.id = oc.idb = 1.0;
    if constexpr (lds_stageable<Src>::value) {
        // the LDS-staged pass where the engine's passes use it (the same rule)
        constexpr int W = TX * VecOf<double>::V;
        if (F.dim == 3 && !F.gal && s->lds != 0 && (F.mx - 1) % W == 0 && (s->lds == 1 || F.mz >= 400))
            return s->cx.launch(cls, true, bytes, [&] {
                k_stencil_lds<double, OpConst<double>, Src, Out, TY, 1><<<g, bk, 0, st>>>(
                    F.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz), oc, src, out, part);
            });
    }
    return s->cx.launch(cls, true, bytes, [&] {
        if (F.dim == 3)
            k_stencil<double, 3, OpConst<double>, Src, Out><<<g, bk, 0, st>>>(F.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz), oc, src, out, part);
        else if (F.dim == 2)
            k_stencil<double, 2, OpConst<double>, Src, Out><<<g, bk, 0, st>>>(F.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz), oc, src, out, part);
        else
            k_stencil<double, 1, OpConst<double>, Src, Out><<<g, bk, 0, st>>>(F.mx, F.my, F.mz, kchf(F.dim, F.my, F.mz), oc, src, out, part);
    });
}

// x += alpha_0 p_0 + ... + alpha_{m-1} p_{m-1} (f.m = m terms, in order) on
// the finest level, then the list is empty
static int fold_x(p4g_solver_s* s, double* x, FoldList& f, bool init) {
    return launch_fold_x(&s->cx, s->lv.back().n, x, f, init);
}

// Mixed precision: FP32 PCG (FP32 vectors, FP32 V-cycle, FP64 coarse
// solve and dot accumulation) with FP64 residual replacement ("reliable
// updates", van der Vorst & Ye, SISC 22 (2000)): when the iterated residual
// has fallen by delta (or below the target), r = b - A x is recomputed in
// FP64 in the pass that applies the iteration's update to x.  The iterate
// update x += alpha p goes to the FP64 iterate (-p4g_mixed_x64 1): the
// directions and step lengths of up to -p4g_x_defer iterations are kept and
// the replacement pass applies them in order (a replacement is forced when
// the list is full), or with -p4g_x_defer 0 the next direction pass applies
// each; or (-p4g_mixed_x64 0) the next direction pass adds it to an FP32
// correction x32 holding the updates since the last replacement, which the
// replacement pass folds into the FP64 iterate.  The Krylov recurrence continues across
// replacements; convergence is decided on that FP64 residual:
// |b - A x|_2 <= max(rtol |b|_2, atol).
static int mixed_solve(p4g_solver_s* s, const double* b, double* x) {
    Engine<float>& e = *s->e32;
    const LevelInfo& F = s->lv.back();
    const long n = F.n;
    hipStream_t st = s->cx.stream;
    Ctx* cx = &s->cx;
    double* S = cx->S;          // engine scalars
    double* So = cx->S + 16;    // outer scalars
    const double* hs = cx->hS;
    const double* ho = cx->hS + 16;
    const int gb = vblocks(n);
    float* x32 = e.xbuf;        // correction since the last replacement (x64 off)
    float* r32 = e.r;
    const bool x64 = s->mixed_x64;    // deferred updates into the FP64 iterate (else into x32)
    const int D = x64 ? mixed_defer(s) : 0;
    // the FP64 iterate: in place in x when it is never swapped (deferred
    // updates applied in place), else in xa / xb and copied to x at the end
    double* xcur = D > 0 && x != b ? x : s->xa;
    double* xnext = s->xb;
    bool xset = D == 0;   // with deferred updates the first fold writes x (x zeroed on an exit before)
    P4G_TRY(cx->launch(P4G_K_OTHER, true, 20.0 * n, [&] {
        if (xset) k_fill<double><<<gb, 256, 0, st>>>(n, 0.0, xcur);
        if (D == 0) k_fill<float><<<gb, 256, 0, st>>>(n, 0.0f, e.pa);   // p_{-1} = 0 (with D: z, beta_0 = 0)
        k_convert_norm<double, float><<<gb, 256, 0, st>>>(n, b, r32, cx->part);   // r32 = b, (b, b)
    }));
    P4G_TRY(cx->finalize(gb, So, FIN_RAW, 0));
    long nb = 0;
    P4G_TRY(e.pc_apply_dots(r32, e.z, &nb));
    P4G_TRY(cx->finalize(nb, S, FIN_BETA_PRECOND, 1));
    P4G_TRY(cx->readS(cx->hS));
    const double bnorm = std::sqrt(ho[S_TMP0]);
    const double target = std::max(s->rtol * bnorm, s->atol);
    p4g_Report& R = s->rep;
    R.rnorm0 = bnorm;
    R.rnorm = bnorm;
    R.inner_its = 0;
    R.its = 0;
    R.reason = 0;
    for (int q = 0; q < 64; ++q) R.history[q] = 0.0;
    R.history[0] = bnorm;
    s->hist.assign(1, bnorm);
    if (bnorm <= s->atol) R.reason = P4G_CONVERGED_ATOL;
    const double delta = s->reliable_delta;
    double rr_last = bnorm * bnorm;   // |r|^2 at the last replacement
    bool replace_now = false;
    bool for_target = false;          // the requested replacement may end the solve (|r| <= target)
    bool x32live = false;             // x32 holds updates not yet folded into the FP64 iterate
    float* pold = D > 0 ? e.z : e.pa;   // p_{-1} (beta_0 = 0)
    float* pnew = e.pb;
    double dpi_old = 0.0;
    double rz_old = hs[S_RZ];         // beta of the previous iteration (its sign is tested)
    int i = 0, nrep = 0;
    float* r32n = e.r2;               // the residual update goes to r32n
    bool xpend = false;               // x += alpha p of the last iteration not yet applied
    // deferred FP64 iterate updates: p_j in a ring of FP32 buffers (never
    // overwritten while listed), alpha_j copied to S[32 + j]
    std::vector<float*> ring = {e.pa, e.pb};
    for (int j = 0; j < D; ++j) ring.push_back(s->pdef[j]);
    FoldList fl{};
    double* Sal = cx->S + 32;
    auto listed = [&](const float* p) {
        for (int j = 0; j < fl.m; ++j)
            if (fl.p[j] == p) return true;
        return false;
    };
    // direction update (with the deferred iterate update when one is
    // pending); the next one is queued before the host waits for the
    // scalars of the current iteration (it needs no host decision)
    auto direction = [&]() -> int {
        long nbs = 0;
        if (D) {   // the new direction into a buffer neither current nor listed
            pnew = nullptr;
            for (float* b : ring)
                if (b != pold && !listed(b)) { pnew = b; break; }
            if (!pnew) return set_error(PETSC_ERR_LIB, "internal: direction buffers exhausted");
        }
        if (xpend && x64) {
            P4G_TRY(e.direction<double>(e.z, pold, pnew, xcur, S + S_BCOEF, S + S_ALPHA, &nbs));
        } else if (xpend) {
            P4G_TRY(e.direction<float>(e.z, pold, pnew, x32, S + S_BCOEF, S + S_ALPHA, &nbs, !x32live));
            x32live = true;
        } else
            P4G_TRY(e.stencil(s->nlev - 1, SrcDir<float>{e.z, pold, pnew, S + S_BCOEF}, OutApplyDot<float>{e.w},
                              P4G_K_STENCIL_DOT, 4, &nbs, false));
        return cx->finalize(nbs, S, FIN_ALPHA, 0);
    };
    if (!R.reason) P4G_TRY(direction());
    for (; i < s->maxit && !R.reason; ++i) {
        R.its = i + 1;
        if (hs[S_RZ] == 0.0) replace_now = true;   // exact breakdown: certify by replacement
        if (D && fl.m == D) replace_now = true;    // the deferred list is full
        std::swap(pold, pnew);
        const bool replaced = replace_now;
        if (replace_now) {
            // x <- x (+ x32) + alpha p and r = b - A x in FP64 in one pass,
            // or the deferred updates and this one applied to x in order and
            // then r = b - A x (the residual pass then stages x in LDS); the
            // recursive residual update of this iteration is not needed
            long nbk = 0;
            if (D) {
                fl.p[fl.m] = pold;
                fl.a[fl.m] = S + S_ALPHA;
                ++fl.m;
                P4G_TRY(fold_x(s, xcur, fl, !xset));
                xset = true;
                P4G_TRY(fine64(s, SrcArr<double>{xcur}, OutResidCvt<double, float>{b, r32}, P4G_K_OTHER, 28.0 * n,
                               &nbk));
            } else {
                if (x32live)
                    P4G_TRY(fine64(s, SrcAxpySumPut<double, float>{xcur, x32, pold, S + S_ALPHA, xnext},
                                   OutResidCvt<double, float>{b, r32}, P4G_K_OTHER, 36.0 * n, &nbk));
                else
                    P4G_TRY(fine64(s, SrcAxpyPut<double, float>{xcur, pold, S + S_ALPHA, xnext},
                                   OutResidCvt<double, float>{b, r32}, P4G_K_OTHER, 32.0 * n, &nbk));
                std::swap(xcur, xnext);
            }
            P4G_TRY(cx->finalize(nbk, So, FIN_RAW, 0));
            xpend = false;
            x32live = false;
            replace_now = false;
            ++nrep;
            P4G_TRY(e.pc_apply_dots(r32, e.z, &nb));
        } else {
            // r - alpha w fused into the first kernel of the V-cycle
            e.pend.r = r32;
            e.pend.w = e.w;
            e.pend.alpha = S + S_ALPHA;
            e.pend.rn = r32n;
            e.pend.rr = true;
            e.pend.on = true;
            if (D) {   // x += alpha p deferred to the next replacement
                P4G_TRY(cx->launch(P4G_K_OTHER, false, 0.0, [&] {
                    k_copy_scalar<<<1, 64, 0, st>>>(S + S_ALPHA, Sal + fl.m);
                }));
                fl.p[fl.m] = pold;
                fl.a[fl.m] = Sal + fl.m;
                ++fl.m;
            } else {
                xpend = true;
            }
            P4G_TRY(e.pc_apply_dots(r32n, e.z, &nb));
            std::swap(r32, r32n);
            P4G_TRY(cx->finalize(e.pend.nparts, S, FIN_RR, 0, cx->part_rr));
        }
        P4G_TRY(cx->finalize(nb, S, FIN_BETA_PRECOND, 0));
        // the next direction is queued before the host reads the scalars,
        // unless this replacement was asked for by an iterated residual
        // below the target: then the host tests first, and a converged
        // solve spends no pass on a direction it does not use (a
        // replacement iteration leaves no pending x update to apply)
        const bool spec = !(replaced && for_target);
        P4G_TRY(cx->readS_begin(cx->hS));
        if (spec) P4G_TRY(direction());   // applies a pending x += alpha p: x is complete on every exit
        P4G_TRY(cx->readS_end());
        const double dpi = hs[S_PW];      // (p, A p) of this iteration's direction (read before a re-read)
        const bool indefinite = dpi == 0.0 || (i > 0 && ((dpi > 0) != (dpi_old > 0)));
        if (!spec && !indefinite) {
            const double rt = std::sqrt(ho[S_TMP0]);
            if (rt <= target) {
                R.rnorm = rt;
                if (i + 1 < 64) R.history[i + 1] = rt;
                s->hist.push_back(rt);
                R.reason = rt < s->atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL;
                break;
            }
            P4G_TRY(direction());
            P4G_TRY(cx->readS(cx->hS));
        }
        for_target = false;
        if (indefinite) { R.reason = P4G_DIVERGED_INDEFINITE_MAT; break; }
        dpi_old = dpi;
        double rn;
        if (replaced) {
            rn = std::sqrt(ho[S_TMP0]);
            rr_last = rn * rn;
        } else {
            rn = std::sqrt(hs[S_RR]);
        }
        R.rnorm = rn;
        if (i + 1 < 64) R.history[i + 1] = rn;
        s->hist.push_back(rn);
        if (std::isnan(rn) || std::isinf(rn)) { R.reason = P4G_DIVERGED_NANORINF; break; }
        if (replaced && rn <= target) { R.reason = rn < s->atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL; break; }
        if (rn >= s->dtol * bnorm) { R.reason = P4G_DIVERGED_DTOL; break; }
        if (!replaced && (rn <= target || rn * rn <= delta * delta * rr_last)) {
            replace_now = true;
            for_target = rn <= target;
        }
        // beta changing sign (an indefinite V-cycle): stopped as PETSc's CG does
        const double rz = hs[S_RZ];
        if (rz * rz_old < 0.0 && i + 1 < s->maxit) {
            R.its = i + 2;
            R.reason = P4G_DIVERGED_INDEFINITE_PC;
            break;
        }
        rz_old = rz;
    }
    if (!R.reason) R.reason = P4G_DIVERGED_ITS;
    R.inner_its = nrep;   // number of FP64 residual replacements
    if (fl.m) {           // deferred updates not yet applied
        P4G_TRY(fold_x(s, xcur, fl, !xset));
        xset = true;
    }
    if (!xset) {          // no update applied: x = 0
        P4G_TRY(cx->launch(P4G_K_OTHER, true, 8.0 * n, [&] { k_fill<double><<<gb, 256, 0, st>>>(n, 0.0, xcur); }));
    }
    if (x32live) {        // fold the last correction: x = x + x32
        P4G_TRY(cx->launch(P4G_K_OTHER, true, 20.0 * n, [&] {
            k_axpy_mixed<<<gb, 256, 0, st>>>(n, xcur, x32);
        }));
    }
    if (xcur != x) HIP_TRY(hipMemcpyAsync(x, xcur, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return 0;
}

// |b - A x|_2 / |b|_2 in FP64 on the finest level (one pass, nothing stored)
static int true_residual(p4g_solver_s* s, const double* b, const double* x, double* rel) {
    const long n = s->lv.back().n;
    double* S = s->cx.S + 16;
    long nbk = 0;
    // (b - A x, b - A x) and (b, b) in one pass, nothing stored
    P4G_TRY(fine64(s, SrcArr<double>{x}, OutResidNorms<double>{b}, P4G_K_OTHER, 16.0 * n, &nbk));
    P4G_TRY(s->cx.finalize(nbk, S, FIN_RAW, 0));
    P4G_TRY(s->cx.readS(s->cx.hS));
    const double rr = s->cx.hS[16 + S_TMP0], bb = s->cx.hS[16 + S_TMP1];
    *rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    return 0;
}

// a level operator in PETSc's binary AIJ format: class id, rows, columns,
// nonzeros, the row lengths, the column indices (ascending per row) and
// the values, written in three passes over the rows
static int mat_view_binary(const LevelInfo& L, const char* path, bool append) {
    if (L.n > 2147483647L) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "matrix size exceeds int32");
    const long sxy = (long)L.mx * L.my;
    // entries of row q in ascending column order: -s[4] (k-1), -s[2] (j-1),
    // -s[0] (i-1), the diagonal, -s[1], -s[3], -s[5]
    auto row = [&](long q, long* col, double* val) -> int {
        const int i = (int)(q % L.mx), j = (int)((q / L.mx) % L.my), k = (int)(q / sxy);
        double dg, c[6];
        host_row(L, i, j, k, dg, c);
        const long off[7] = {-sxy, -(long)L.mx, -1, 0, 1, (long)L.mx, sxy};
        const double v[7] = {-c[4], -c[2], -c[0], dg, -c[1], -c[3], -c[5]};
        int m = 0;
        for (int e = 0; e < 7; ++e)
            if (e == 3 || v[e] != 0.0) {
                col[m] = q + off[e];
                val[m] = v[e];
                ++m;
            }
        return m;
    };
    long nz = 0, col[7];
    double val[7];
    for (long q = 0; q < L.n; ++q) nz += row(q, col, val);
    if (nz > 2147483647L) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "nonzero count exceeds int32");
    BinaryFile f;
    if (!f.open(path, append ? "ab" : "wb")) return set_error(PETSC_ERR_FILE_OPEN, "cannot open %s", path);
    f.i32(kMatClassId);
    f.i32((int32_t)L.n);
    f.i32((int32_t)L.n);
    f.i32((int32_t)nz);
    for (long q = 0; q < L.n; ++q) f.i32(row(q, col, val));
    for (long q = 0; q < L.n; ++q) {
        const int m = row(q, col, val);
        for (int e = 0; e < m; ++e) f.i32((int32_t)col[e]);
    }
    for (long q = 0; q < L.n; ++q) {
        const int m = row(q, col, val);
        for (int e = 0; e < m; ++e) f.f64(val[e]);
    }
    if (!f.close()) return set_error(PETSC_ERR_FILE_WRITE, "write to %s failed", path);
    return 0;
}

// KSPSolve's viewers: -ksp_view_mat, -ksp_view_rhs, -ksp_view_solution
// binary:<file>, after a solve (b, x on the device); a file named by
// several of them receives the objects in that order
static int ksp_views(p4g_solver_s* s, const double* b, double* x) {
    static const char* names[3] = {"ksp_view_mat", "ksp_view_rhs", "ksp_view_solution"};
    std::vector<std::string> written;
    std::vector<double> h;
    for (int w = 0; w < 3; ++w) {
        if (!s->opt.has(names[w])) continue;
        const std::string path = s->opt.gets(names[w], "").substr(7);   // after "binary:"
        bool append = false;
        for (const std::string& p : written) append = append || p == path;
        written.push_back(path);
        if (w == 0) {
            P4G_TRY(mat_view_binary(s->lv.back(), path.c_str(), append));
            continue;
        }
        const long n = s->lv.back().n;
        h.resize(n);
        HIP_TRY(hipMemcpy(h.data(), w == 1 ? b : x, n * sizeof(double), hipMemcpyDeviceToHost));
        P4G_TRY(p4g_vec_view_binary(path.c_str(), n, h.data(), append ? 1 : 0));
    }
    return 0;
}

static int sor_pipes_reset(p4g_solver_s* s);
static int sor_pipes_check(p4g_solver_s* s, const char* who);

int ksp_solve_device(p4g_solver_s* s, const double* b, double* x) {
    P4G_TRY(solver_setup(s));
    P4G_TRY(sor_pipes_reset(s));
    hipStream_t st = s->cx.stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    p4g_Report& R = s->rep;
    if (s->mixed) {
        P4G_TRY(mixed_solve(s, b, x));
    } else {
        typename Engine<double>::CGOut co;
        if (s->ksp_rich)
            P4G_TRY(s->e64->richardson(b, x, s->normtype, s->rtol, s->atol, s->dtol, s->maxit, s->rich_outer_scale,
                                       s->monitor, co, s->cx.S));
        else if (s->ksp_gmres)
            P4G_TRY(s->e64->gmres(b, x, s->rtol, s->atol, s->dtol, s->maxit, s->gmres_restart, co, s->cx.S));
        else
            P4G_TRY(s->e64->cg(b, x, s->normtype, s->rtol, s->atol, s->dtol, s->maxit, co, s->cx.S));
        R.its = co.its; R.reason = co.reason; R.rnorm0 = co.rnorm0; R.rnorm = co.rnorm;
        R.inner_its = 0;
        for (size_t q = 0; q < 64; ++q) R.history[q] = q < co.hist.size() ? co.hist[q] : 0.0;
        s->hist = co.hist;
    }
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    R.solve_ms = ms;
    s->cx.tm.collect();
    P4G_TRY(sor_pipes_check(s, "KSPSolve"));
    double rel = 0.0;
    P4G_TRY(true_residual(s, b, x, &rel));
    R.true_relres = rel;
    return ksp_views(s, b, x);
}

}  // namespace p4g

// ================================================================== C ABI

extern "C" {

int p4g_da_create(int dim, int mx, int my, int mz, double Lx, double Ly, double Lz, DM* da) {
    clear_error();
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_create: null output");
    if (dim < 1 || dim > 3) return set_error(1, "invalid dim for DMDA creation\n");
    if (mx < 2 || (dim > 1 && my < 2) || (dim > 2 && mz < 2))
        return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_da_create: need at least 2 points per dimension");
    if (!(Lx > 0) || (dim > 1 && !(Ly > 0)) || (dim > 2 && !(Lz > 0)))
        return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_da_create: domain lengths must be positive");
    p4g_da_s* d = new p4g_da_s;
    d->dim = dim; d->mx = mx; d->my = dim > 1 ? my : 1; d->mz = dim > 2 ? mz : 1;
    d->L[0] = Lx; d->L[1] = dim > 1 ? Ly : 1.0; d->L[2] = dim > 2 ? Lz : 1.0;
    *da = d;
    return 0;
}

int p4g_da_set_from_options(DM da, const char* options) {
    clear_error();
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_set_from_options: null DM");
    Options o;
    o.parse(options);
    int m[3] = {o.geti("da_grid_x", da->mx), o.geti("da_grid_y", da->my), o.geti("da_grid_z", da->mz)};
    const int r = o.geti("da_refine", 0);
    if (r < 0 || r > 14) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "-da_refine %d out of range", r);
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < r; ++q) m[d] = 2 * (m[d] - 1) + 1;
    da->mx = m[0];
    da->my = da->dim > 1 ? m[1] : 1;
    da->mz = da->dim > 2 ? m[2] : 1;
    da->refine = r;
    return 0;
}

int p4g_da_get_sizes(DM da, int* dim, int* mx, int* my, int* mz, int* refine) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_get_sizes: null DM");
    if (dim) *dim = da->dim;
    if (mx) *mx = da->mx;
    if (my) *my = da->my;
    if (mz) *mz = da->mz;
    if (refine) *refine = da->refine;
    return 0;
}

int p4g_da_set_function_local(DM da, DMDASNESFunctionFn* fn, void* ctx) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_set_function_local: null DM");
    da->fn = fn;
    da->fnctx = ctx;
    return 0;
}

int p4g_da_set_jacobian_local(DM da, DMDASNESJacobianFn* fn, void* ctx) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_set_jacobian_local: null DM");
    da->jac = fn;
    da->jacctx = ctx;
    return 0;
}

int p4g_da_set_cell_coefficient(DM da, const double* kcell) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_set_cell_coefficient: null DM");
    if (da->dim != 3) return set_error(PETSC_ERR_SUP, "cell coefficients need a 3D grid");
    const size_t nc = (size_t)(da->mx - 1) * (da->my - 1) * (da->mz - 1);
    if (!kcell) { da->kcell.clear(); return 0; }
    for (size_t q = 0; q < nc; ++q)
        if (!(kcell[q] > 0.0)) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "cell coefficient must be positive");
    da->kcell.assign(kcell, kcell + nc);
    return 0;
}

int p4g_da_destroy(DM* da) {
    if (da && *da) { delete *da; *da = nullptr; }
    return 0;
}

int p4g_solver_create(DM da, const char* options, p4g_Solver* out) {
    clear_error();
    if (!da || !out) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_create: null argument");
    p4g_solver_s* s = new p4g_solver_s;
    s->da = da;
    s->opts_str = options ? options : "";
    s->opt.parse(options);
    std::memset(&s->rep, 0, sizeof s->rep);
    int e = configure(s);
    if (e) { delete s; return e; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete s;
        return set_error(PETSC_ERR_LIB, "p4g_solver_create: no HIP device is available (the solver "
                                        "runs on the GPU only)");
    }
    *out = s;
    return 0;
}

int p4g_solver_setup(p4g_Solver s) {
    clear_error();
    if (!s) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_setup: null solver");
    return solver_setup(s);
}

int p4g_da_mat_view_binary(DM da, const char* path, int append) {
    clear_error();
    if (!da || !path) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_mat_view_binary: null argument");
    p4g_solver_s tmp;   // one level: the rows of the callbacks on this grid (no device work)
    tmp.da = da;
    tmp.nlev = 1;
    P4G_TRY(build_levels(&tmp));
    return mat_view_binary(tmp.lv.back(), path, append != 0);
}

int p4g_da_mat_is_symmetric(DM da, const double* u, double tol, int* symmetric) {
    clear_error();
    if (!da || !symmetric) return set_error(PETSC_ERR_ARG_WRONG, "p4g_da_mat_is_symmetric: null argument");
    if (!da->jac) return set_error(PETSC_ERR_ARG_WRONG, "no Jacobian callback: call p4g_da_set_jacobian_local");
    const long n = da->n();
    p4g_mat_s M;
    M.da = da;
    M.seen.assign(n, 0);
    M.keep = true;
    std::vector<double> ux(u ? u : nullptr, u ? u + n : nullptr);
    if (!u) ux.assign(n, 0.0);
    std::vector<double*> rows;
    std::vector<double**> planes;
    void* au = make_view(da, ux.data(), rows, planes);
    DMDALocalInfo info;
    info_of(da, &info);
    P4G_TRY(da->jac(&info, au, &M, &M, da->jacctx));
    // INSERT_VALUES: the last value inserted at (row, column) is the entry
    const size_t ne = M.ev.size();
    std::vector<size_t> ord(ne);
    for (size_t q = 0; q < ne; ++q) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return M.er[a] != M.er[b] ? M.er[a] < M.er[b] : M.ec[a] < M.ec[b];
    });
    std::vector<long> r, c;
    std::vector<double> v;
    for (size_t q = 0; q < ne; ++q) {
        const size_t e = ord[q];
        if (!r.empty() && r.back() == M.er[e] && c.back() == M.ec[e]) { v.back() = M.ev[e]; continue; }
        r.push_back(M.er[e]); c.push_back(M.ec[e]); v.push_back(M.ev[e]);
    }
    // |a_ij - a_ji| <= tol for every entry (an absent entry is zero)
    auto find = [&](long i, long j) -> double {
        size_t lo = 0, hi = r.size();
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (r[mid] < i || (r[mid] == i && c[mid] < j)) lo = mid + 1; else hi = mid;
        }
        return (lo < r.size() && r[lo] == i && c[lo] == j) ? v[lo] : 0.0;
    };
    int sym = 1;
    for (size_t q = 0; q < r.size() && sym; ++q)
        if (!(std::fabs(v[q] - find(c[q], r[q])) <= tol)) sym = 0;
    *symmetric = sym;
    return 0;
}

int p4g_ksp_solve_device(p4g_Solver s, const double* d_b, double* d_x, p4g_Report* rep) {
    clear_error();
    if (!s || !d_b || !d_x) return set_error(PETSC_ERR_ARG_WRONG, "p4g_ksp_solve_device: null argument");
    P4G_TRY(ksp_solve_device(s, d_b, d_x));
    if (rep) *rep = s->rep;
    return 0;
}

int p4g_ksp_solve(p4g_Solver s, const double* b, double* x, p4g_Report* rep) {
    clear_error();
    if (!s || !b || !x) return set_error(PETSC_ERR_ARG_WRONG, "p4g_ksp_solve: null argument");
    P4G_TRY(solver_setup(s));
    const long n = s->lv.back().n;
    double *db = nullptr, *dx = nullptr;
    HIP_TRY(hipMalloc(&db, n * sizeof(double)));
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess) {
        hipFree(db);
        return set_error(PETSC_ERR_MEM, "p4g_ksp_solve: hipMalloc failed");
    }
    int e = 0;
    if (hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        e = set_error(PETSC_ERR_LIB, "p4g_ksp_solve: upload failed");
    if (!e) e = ksp_solve_device(s, db, dx);
    if (!e && hipMemcpy(x, dx, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        e = set_error(PETSC_ERR_LIB, "p4g_ksp_solve: download failed");
    hipFree(db);
    hipFree(dx);
    if (!e && rep) *rep = s->rep;
    return e;
}

int p4g_snes_solve(p4g_Solver s, double* u, p4g_Report* rep) {
    clear_error();
    if (!s || !u) return set_error(PETSC_ERR_ARG_WRONG, "p4g_snes_solve: null argument");
    p4g_da_s* da = s->da;
    if (!da->fn) return set_error(PETSC_ERR_ARG_WRONG, "no residual callback: call p4g_da_set_function_local");
    const long n = da->n();
    std::vector<double> F(n, 0.0), y(n, 0.0);
    std::vector<double*> r1, r2;
    std::vector<double**> p1, p2;
    void* au = make_view(da, u, r1, p1);
    void* aF = make_view(da, F.data(), r2, p2);
    DMDALocalInfo info;
    info_of(da, &info);
    P4G_TRY(da->fn(&info, au, aF, da->fnctx));   // F(u0)
    double f0 = 0.0;
    for (double v : F) f0 += v * v;
    P4G_TRY(p4g_ksp_solve(s, F.data(), y.data(), nullptr));
    for (long q = 0; q < n; ++q) u[q] -= y[q];
    P4G_TRY(da->fn(&info, au, aF, da->fnctx));   // F(u)
    double f1 = 0.0;
    for (double v : F) f1 += v * v;
    s->rep.fnorm0 = std::sqrt(f0);
    s->rep.fnorm1 = std::sqrt(f1);
    if (rep) *rep = s->rep;
    return 0;
}

int p4g_pc_apply_device(p4g_Solver s, const double* d_r, double* d_z) {
    clear_error();
    if (!s || !d_r || !d_z) return set_error(PETSC_ERR_ARG_WRONG, "p4g_pc_apply_device: null argument");
    P4G_TRY(solver_setup(s));
    if (s->mixed) {
        Engine<float>& e = *s->e32;
        const long n = s->lv.back().n;
        hipStream_t st = s->cx.stream;
        k_convert<double, float><<<vblocks(n), 256, 0, st>>>(n, d_r, s->r32);
        P4G_TRY(e.pc_apply(s->r32, s->e32v));
        k_convert<float, double><<<vblocks(n), 256, 0, st>>>(n, s->e32v, d_z);
    } else {
        P4G_TRY(s->e64->pc_apply(d_r, d_z));
    }
    HIP_TRY(hipStreamSynchronize(s->cx.stream));
    s->cx.tm.collect();
    return 0;
}

// ---------------------------------------------- no-allocation PCApply
// The work arrays of one multigrid cycle of p4g_pc_apply_device_noalloc are
// carved from this file-scope device array instead of the solver's own
// allocations.  Every GPU holds its own copy (at its own address, looked up
// once per device with hipGetSymbolAddress) until the process ends; calls on
// one GPU share it and overlapping calls race.  The cycle writes every work
// element before it reads it, so nothing in it is cleared between calls.
constexpr size_t kNoAllocBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kNoAllocDevices = 64;
static std::mutex noalloc_mu;
static unsigned char* noalloc_ptr[kNoAllocDevices];

static int noalloc_scratch(unsigned char** base) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices)
        return set_error(PETSC_ERR_SUP, "p4g_pc_apply_device_noalloc: device %d beyond the %d supported", dev,
                         kNoAllocDevices);
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(PETSC_ERR_LIB, "p4g_pc_apply_device_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                             hipGetErrorString(e));
        noalloc_ptr[dev] = (unsigned char*)p;
    }
    *base = noalloc_ptr[dev];
    return 0;
}

int p4g_pc_apply_device_noalloc(p4g_Solver s, const double* d_r, double* d_z) {
    clear_error();
    if (!s || !d_r || !d_z) return set_error(PETSC_ERR_ARG_WRONG, "p4g_pc_apply_device_noalloc: null argument");
    if (!s->ready)
        return set_error(PETSC_ERR_ORDER, "p4g_pc_apply_device_noalloc: set the solver up first (p4g_solver_setup); "
                                          "this entry point allocates nothing");
    if (s->mixed || !s->e64)
        return set_error(PETSC_ERR_SUP, "p4g_pc_apply_device_noalloc: FP64 solvers only (-p4g_precision double)");
    Engine<double>& E = *s->e64;
    if (E.pend.on) return set_error(PETSC_ERR_LIB, "internal: pending residual update");
    // the carve: each work array at a 256-byte aligned offset, then the
    // block partial sums (two buffers of 2 npart_cap doubles)
    size_t off = 0;
    auto room = [&](long ndoubles) -> size_t {
        const size_t o = off;
        off += ((size_t)ndoubles * sizeof(double) + 255) & ~size_t(255);
        return o;
    };
    E.each_work([&](double*& p, long n) { if (p) room(n); });
    room(2 * s->cx.npart_cap);
    room(2 * s->cx.npart_cap);
    if (off > kNoAllocBytes)
        return set_error(PETSC_ERR_ARG_OUTOFRANGE,
                         "p4g_pc_apply_device_noalloc: this hierarchy's cycle needs %zu bytes of work arrays, more than "
                         "the %zu MiB scratch (FP64 V(1,1) hierarchies up to 129^3 or 2049^2 fit)", off,
                         kNoAllocBytes >> 20);
    unsigned char* base = nullptr;
    P4G_TRY(noalloc_scratch(&base));
    // swap the carve in, run the cycle, swap the solver's arrays back
    constexpr int kMaxWork = 8 * 32;
    double* saved[kMaxWork];
    int nsaved = 0;
    E.each_work([&](double*&, long) { ++nsaved; });
    if (nsaved > kMaxWork) return set_error(PETSC_ERR_SUP, "p4g_pc_apply_device_noalloc: too many levels");
    off = 0;
    nsaved = 0;
    E.each_work([&](double*& p, long n) {
        saved[nsaved++] = p;
        if (p) p = (double*)(base + room(n));
    });
    double* const part0 = s->cx.part;
    double* const part_rr0 = s->cx.part_rr;
    s->cx.part = (double*)(base + room(2 * s->cx.npart_cap));
    s->cx.part_rr = (double*)(base + room(2 * s->cx.npart_cap));
    int on[P4G_K_COUNT];   // HIP-event timing off: its event pool grows by allocation
    for (int c = 0; c < P4G_K_COUNT; ++c) { on[c] = s->cx.tm.on[c]; s->cx.tm.on[c] = 0; }
    int e = E.pc_apply(d_r, d_z);
    if (!e) {
        const hipError_t h = hipStreamSynchronize(s->cx.stream);
        if (h != hipSuccess) e = set_error(PETSC_ERR_LIB, "p4g_pc_apply_device_noalloc: %s", hipGetErrorString(h));
    }
    size_t q = 0;
    E.each_work([&](double*& p, long) { p = saved[q++]; });
    s->cx.part = part0;
    s->cx.part_rr = part_rr0;
    for (int c = 0; c < P4G_K_COUNT; ++c) s->cx.tm.on[c] = on[c];
    return e;
}

int p4g_pc_apply(p4g_Solver s, const double* r, double* z) {
    clear_error();
    if (!s || !r || !z) return set_error(PETSC_ERR_ARG_WRONG, "p4g_pc_apply: null argument");
    P4G_TRY(solver_setup(s));
    const long n = s->lv.back().n;
    double *dr = nullptr, *dz = nullptr;
    HIP_TRY(hipMalloc(&dr, n * sizeof(double)));
    if (hipMalloc(&dz, n * sizeof(double)) != hipSuccess) { hipFree(dr); return set_error(PETSC_ERR_MEM, "hipMalloc"); }
    int e = 0;
    if (hipMemcpy(dr, r, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) e = set_error(PETSC_ERR_LIB, "upload");
    if (!e) e = p4g_pc_apply_device(s, dr, dz);
    if (!e && hipMemcpy(z, dz, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) e = set_error(PETSC_ERR_LIB, "download");
    hipFree(dr);
    hipFree(dz);
    return e;
}

int p4g_mat_mult_device(p4g_Solver s, const double* d_x, double* d_y) {
    clear_error();
    if (!s || !d_x || !d_y) return set_error(PETSC_ERR_ARG_WRONG, "p4g_mat_mult_device: null argument");
    P4G_TRY(solver_setup(s));
    P4G_TRY(fine64(s, SrcArr<double>{d_x}, OutApply<double>{d_y}, P4G_K_APPLY, 16.0 * s->lv.back().n, nullptr));
    s->cx.tm.collect();
    HIP_TRY(hipStreamSynchronize(s->cx.stream));
    return 0;
}

int p4g_mat_mult(p4g_Solver s, const double* x, double* y) {
    clear_error();
    if (!s || !x || !y) return set_error(PETSC_ERR_ARG_WRONG, "p4g_mat_mult: null argument");
    P4G_TRY(solver_setup(s));
    const long n = s->lv.back().n;
    double *dx = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc(&dx, n * sizeof(double)));
    if (hipMalloc(&dy, n * sizeof(double)) != hipSuccess) { hipFree(dx); return set_error(PETSC_ERR_MEM, "hipMalloc"); }
    int e = 0;
    if (hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) e = set_error(PETSC_ERR_LIB, "upload");
    if (!e) e = p4g_mat_mult_device(s, dx, dy);
    if (!e && hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) e = set_error(PETSC_ERR_LIB, "download");
    hipFree(dx);
    hipFree(dy);
    return e;
}

int p4g_solver_get_history(p4g_Solver s, double* h, int cap, int* n) {
    if (!s || !n || (cap > 0 && !h)) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_get_history: bad argument");
    *n = (int)s->hist.size();
    for (int q = 0; q < std::min(cap, *n); ++q) h[q] = s->hist[q];
    return 0;
}

int p4g_solver_get_report(p4g_Solver s, p4g_Report* rep) {
    if (!s || !rep) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_get_report: null argument");
    *rep = s->rep;
    return 0;
}

int p4g_solver_destroy(p4g_Solver* s) {
    if (s && *s) {
        if ((*s)->cx.stream) hipStreamSynchronize((*s)->cx.stream);
        delete *s;
        *s = nullptr;
    }
    return 0;
}

int p4g_solver_set_timing(p4g_Solver s, int cls, int on) {
    if (!s || cls < 0 || cls >= P4G_K_COUNT) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_set_timing");
    if (on < 0 || on > 2) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_solver_set_timing: on must be 0, 1 or 2");
    s->cx.tm.on[cls] = on;
    return 0;
}

int p4g_solver_kernel_stats(p4g_Solver s, int cls, int fine_only, long* launches, double* ms, double* bytes) {
    if (!s || cls < 0 || cls >= P4G_K_COUNT) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_kernel_stats");
    const KStat& k = s->cx.tm.st[cls];
    if (launches) *launches = fine_only ? k.fine_launches : k.launches;
    if (ms) *ms = fine_only ? k.fine_ms : k.ms;
    if (bytes) *bytes = fine_only ? k.fine_bytes : k.bytes;
    return 0;
}

int p4g_solver_reset_stats(p4g_Solver s) {
    if (!s) return set_error(PETSC_ERR_ARG_WRONG, "p4g_solver_reset_stats");
    s->cx.tm.collect();
    for (auto& k : s->cx.tm.st) k = KStat();
    return 0;
}

int p4g_device_count(int* n) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (n) *n = c;
    return 0;
}

int p4g_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

void* p4g_device_malloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error(PETSC_ERR_MEM, "p4g_device_malloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

int p4g_device_free(void* p) {
    if (p) HIP_TRY(hipFree(p));
    return 0;
}

// Copies on a library-owned non-blocking stream, synchronized before return:
// never on the legacy stream, which would implicitly depend on (and, while a
// server thread captures a graph, fail against) the other streams.  They are
// ordered after everything the library's calls have completed.
static int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    static hipStream_t cs = [] {
        hipStream_t t = nullptr;
        if (hipStreamCreateWithFlags(&t, hipStreamNonBlocking) != hipSuccess) t = nullptr;
        return t;
    }();
    if (!cs) return set_error(PETSC_ERR_LIB, "p4g_memcpy: no copy stream");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, cs));
    HIP_TRY(hipStreamSynchronize(cs));
    return 0;
}

int p4g_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}

int p4g_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}

int p4g_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

const char* p4g_version(void) { return "p4g 0.1 (gfx950)"; }

}  // extern "C"

// continuous-batching server (uses the types above)
namespace p4g {
// The stall words of the pipelined SOR sweeps (k_sor9_strips): cleared at the
// start of a solve, read after each linear solve, where the solve has just
// synchronized its stream, so a stall stops the solve at the Newton step it
// happened in.
static int sor_pipes_reset(p4g_solver_s* s) {
    for (SorPipe& P : s->ms_pipes) {
        if (!P.err) continue;
        P.spin_limit = s->sor_test_stall ? 64 : (1L << 24);
        P.need_extra = s->sor_test_stall ? (1ull << 21) : 0;
        HIP_TRY(hipMemsetAsync(P.err, 0, sizeof(int), s->cx.stream));
    }
    return 0;
}

static int sor_pipes_check(p4g_solver_s* s, const char* who) {
    for (const SorPipe& P : s->ms_pipes) {
        if (!P.err) continue;
        int e = 0;
        HIP_TRY(hipMemcpyAsync(&e, P.err, sizeof(int), hipMemcpyDeviceToHost, s->cx.stream));
        HIP_TRY(hipStreamSynchronize(s->cx.stream));
        if (e) return set_error(PETSC_ERR_LIB, "%s: a pipelined SOR sweep stalled", who);
    }
    return 0;
}
}  // namespace p4g

#include "p4g_serve_impl.hpp"
// reduced-space VI Newton (SNESVINEWTONRSLS) over the solver's levels
#include "p4g_vi_impl.hpp"
#include "p4g_ms_impl.hpp"
