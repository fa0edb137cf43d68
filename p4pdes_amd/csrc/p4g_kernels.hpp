// p4g_kernels.hpp -- gfx950 device kernels of the CG + geometric-multigrid
// Poisson path (vertex grids, x fastest, one array per grid function).
//
// Operator rows are those of Poisson{1,2,3}DJacobianLocal
// (c/ch6/poissonfunctions.c:117-258): every row has the diagonal d (boundary
// rows: db), interior rows couple with -sx/-sy/-sz to interior neighbours
// only, boundary rows are diagonal-only.  Q1 transfers act on every vertex,
// boundary included; restriction is the transpose of interpolation.
//
// The stencil kernels march along the slowest dimension (z in 3D, y in 2D):
// a thread owns one (i,j) column, keeps the operand at k-1, k, k+1 in
// registers and reads in-plane neighbours through the cache, so each input
// grid function streams from HBM about once per pass.  One wavefront spans
// 64 consecutive x points starting at i = 1; the lane at i = 1 also owns the
// boundary point i = 0, so (mx-1) = 2^k grids fill every lane.
//
// Reductions are deterministic: each block writes its partial sums, a
// single-block pass adds them in a fixed order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p4g {

constexpr int TX = 64;   // lanes along x: one wavefront
constexpr int TY = 8;    // rows (3D) or march chunks (2D) per block
constexpr int NT = TX * TY;
constexpr int TYF = 6;   // output rows per block of the fused 3D prolongation + smoothing
static_assert(TX * (TYF + 2) == NT, "the fused 3D block (64, TYF+2) has the stencil block's shape");

// Two consecutive grid values, moved as one 2-element vector access.  The
// address need only be aligned to one element (gfx950 global memory takes
// unaligned vector accesses).
template <typename T>
struct T2 {
    T a, b;
};

template <typename T>
__device__ __forceinline__ T2<T> ld2(const T* __restrict__ p) {
    typedef T vec __attribute__((ext_vector_type(2)));
    vec v;
    __builtin_memcpy(&v, p, sizeof(vec));
    return T2<T>{v.x, v.y};
}

template <typename T>
__device__ __forceinline__ void st2(T* __restrict__ p, T a, T b) {
    typedef T vec __attribute__((ext_vector_type(2)));
    vec v;
    v.x = a;
    v.y = b;
    __builtin_memcpy(p, &v, sizeof(vec));
}

// V consecutive grid values moved as one vector access (V * sizeof(T) <= 16).
template <typename T, int V>
struct TV {
    T v[V];
};

template <typename T, int V>
__device__ __forceinline__ TV<T, V> ldv(const T* __restrict__ p) {
    typedef T vec __attribute__((ext_vector_type(V)));
    vec t;
    __builtin_memcpy(&t, p, sizeof(vec));
    TV<T, V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = t[e];
    return r;
}

template <typename T, int V>
__device__ __forceinline__ void stv(T* __restrict__ p, const TV<T, V>& a) {
    typedef T vec __attribute__((ext_vector_type(V)));
    vec t;
#pragma unroll
    for (int e = 0; e < V; ++e) t[e] = a.v[e];
    __builtin_memcpy(p, &t, sizeof(vec));
}

// points per lane of the stencil kernels: one 16-byte access
template <typename T>
struct VecOf {
    static constexpr int V = 16 / sizeof(T);
};

// ------------------------------------------------------------- operators

// Constant-coefficient stencil.
// With diff = 1 an interior row is evaluated as the sum of scaled
// differences sx((uc-w)+(uc-e)) + sy(..) + sz(..), which equals
// d uc - sx(w+e) - sy(s+n) - sz(dn+up) when d = 2(sx+sy+sz) (eliminated
// boundary neighbours enter as 0) and rounds far less for smooth u: the
// differences of close values are exact, where d uc - sum cancels.
template <typename T>
struct OpConst {
    T d, db, sx, sy, sz;   // interior diagonal, boundary diagonal, couplings
    T id, idb;             // Jacobi: 1/d, 1/db (or 1 when no Jacobi scaling)
    int diff = 0;          // 1: difference form (requires d = 2(sx+sy+sz))
    __device__ __forceinline__ void bnd(int, int, int, T u, T& Au, T& idg) const {
        Au = db * u; idg = idb;
    }
    __device__ __forceinline__ void interior(int, int, int, T uc, T w, T e, T s, T n, T dn,
                                             T up, T& Au, T& idg) const {
        if (diff) Au = sx * ((uc - w) + (uc - e)) + sy * ((uc - s) + (uc - n)) + sz * ((uc - dn) + (uc - up));
        else Au = d * uc - sx * (w + e) - sy * (s + n) - sz * (dn + up);
        idg = id;
    }
};

// Variable coefficient -div(k grad u) with cell-centred k (3D).  The face
// coefficient is the mean of the 4 cells sharing the face's edge; a
// boundary row's diagonal is 2(ax+ay+az) times the mean of its cells.
template <typename T>
struct OpVar {
    const T* __restrict__ kc;   // (mx-1)(my-1)(mz-1) cells, x fastest
    int cx, cy, cz;             // cell counts
    T ax, ay, az;               // dvol/h^2 per direction
    int jac;                    // 1: Jacobi scaling, 0: identity
    __device__ __forceinline__ T cell(int i, int j, int k) const {
        return kc[((long)k * cy + j) * cx + i];
    }
    __device__ __forceinline__ void bnd(int i, int j, int k, T u, T& Au, T& idg) const {
        T s = T(0);
        int cnt = 0;
        for (int dk = -1; dk <= 0; ++dk)
            for (int dj = -1; dj <= 0; ++dj)
                for (int di = -1; di <= 0; ++di) {
                    int a = i + di, b = j + dj, c = k + dk;
                    if (a < 0 || b < 0 || c < 0 || a >= cx || b >= cy || c >= cz) continue;
                    s += cell(a, b, c);
                    ++cnt;
                }
        T dg = T(2) * (ax + ay + az) * (s / T(cnt));
        Au = dg * u;
        idg = jac ? T(1) / dg : T(1);
    }
    __device__ __forceinline__ void interior(int i, int j, int k, T uc, T w, T e, T s, T n,
                                             T dn, T up, T& Au, T& idg) const {
        const T c000 = cell(i - 1, j - 1, k - 1), c100 = cell(i, j - 1, k - 1);
        const T c010 = cell(i - 1, j, k - 1), c110 = cell(i, j, k - 1);
        const T c001 = cell(i - 1, j - 1, k), c101 = cell(i, j - 1, k);
        const T c011 = cell(i - 1, j, k), c111 = cell(i, j, k);
        const T q = T(0.25);
        const T kw = q * (c000 + c010 + c001 + c011) * ax;
        const T ke = q * (c100 + c110 + c101 + c111) * ax;
        const T ks = q * (c000 + c100 + c001 + c101) * ay;
        const T kn = q * (c010 + c110 + c011 + c111) * ay;
        const T kd = q * (c000 + c100 + c010 + c110) * az;
        const T ku = q * (c001 + c101 + c011 + c111) * az;
        const T dg = kw + ke + ks + kn + kd + ku;
        Au = dg * uc - kw * w - ke * e - ks * s - kn * n - kd * dn - ku * up;
        idg = jac ? T(1) / dg : T(1);
    }
};

// ----------------------------------------------------- operand sources
//
// A source gives the stencil operand u at a point (load) or at V
// consecutive points (loadv, one vector access), and may store the value
// it formed at the kernel's own points (put / putv).

template <typename T>
struct SrcArr {                       // u = x
    const T* __restrict__ x;
    __device__ __forceinline__ T load(long q) const { return x[q]; }
    template <int V>
    __device__ __forceinline__ TV<T, V> loadv(long q) const { return ldv<T, V>(x + q); }
    __device__ __forceinline__ void put(long, T) const {}
    template <int V>
    __device__ __forceinline__ void putv(long, const TV<T, V>&) const {}
    // LDS staging (k_stencil_lds): the arrays read at stencil neighbours, the
    // operand from their staged values (s: lds_scalar()), the own-point store
    static constexpr int NA = 1;
    __device__ __forceinline__ const T* lds_arr(int) const { return x; }
    __device__ __forceinline__ T lds_scalar() const { return T(0); }
    __device__ __forceinline__ T lds_comb(T v0, T, T) const { return v0; }
    template <int V>
    __device__ __forceinline__ void lds_put(long, const TV<T, V>&, const TV<T, V>&) const {}
};

template <typename T>
struct SrcZero {                      // u = 0 (first smoothing step from a zero guess)
    __device__ __forceinline__ T load(long) const { return T(0); }
    template <int V>
    __device__ __forceinline__ TV<T, V> loadv(long) const {
        TV<T, V> r;
#pragma unroll
        for (int e = 0; e < V; ++e) r.v[e] = T(0);
        return r;
    }
    __device__ __forceinline__ void put(long, T) const {}
    template <int V>
    __device__ __forceinline__ void putv(long, const TV<T, V>&) const {}
};

template <typename T>
struct SrcDir {                       // p_new = z + beta p_old, stored (CG direction)
    const T* __restrict__ z;
    const T* __restrict__ pold;
    T* __restrict__ pnew;
    const double* __restrict__ beta;  // device scalar
    __device__ __forceinline__ T load(long q) const { return z[q] + T(*beta) * pold[q]; }
    template <int V>
    __device__ __forceinline__ TV<T, V> loadv(long q) const {
        const T bt = T(*beta);
        TV<T, V> r = ldv<T, V>(z + q);
        const TV<T, V> pp = ldv<T, V>(pold + q);
#pragma unroll
        for (int e = 0; e < V; ++e) r.v[e] += bt * pp.v[e];
        return r;
    }
    __device__ __forceinline__ void put(long q, T v) const { pnew[q] = v; }
    template <int V>
    __device__ __forceinline__ void putv(long q, const TV<T, V>& u) const { stv<T, V>(pnew + q, u); }
    static constexpr int NA = 2;
    __device__ __forceinline__ const T* lds_arr(int a) const { return a ? pold : z; }
    __device__ __forceinline__ T lds_scalar() const { return T(*beta); }
    __device__ __forceinline__ T lds_comb(T v0, T v1, T bt) const { return v0 + bt * v1; }
    template <int V>
    __device__ __forceinline__ void lds_put(long q, const TV<T, V>& u, const TV<T, V>&) const { stv<T, V>(pnew + q, u); }
};

// p_new = z + beta p_old, stored; at the own point also the deferred CG
// iterate update x += alpha p_old (alpha of the previous iteration; X may be
// wider than T), or x = alpha p_old with init (x not read).
template <typename T, typename X>
struct SrcDirX {
    const T* __restrict__ z;
    const T* __restrict__ pold;
    T* __restrict__ pnew;
    const double* __restrict__ beta;  // device scalar
    X* __restrict__ x;
    const double* __restrict__ alpha; // device scalar
    bool init = false;
    template <int VX>
    __device__ __forceinline__ TV<X, VX> xload(long q) const {
        if (init) {
            TV<X, VX> z0;
#pragma unroll
            for (int e = 0; e < VX; ++e) z0.v[e] = X(0);
            return z0;
        }
        return ldv<X, VX>(x + q);
    }
    __device__ __forceinline__ T load(long q) const { return z[q] + T(*beta) * pold[q]; }
    template <int V>
    __device__ __forceinline__ TV<T, V> loadv(long q) const {
        const T bt = T(*beta);
        TV<T, V> r = ldv<T, V>(z + q);
        const TV<T, V> pp = ldv<T, V>(pold + q);
#pragma unroll
        for (int e = 0; e < V; ++e) r.v[e] += bt * pp.v[e];
        return r;
    }
    __device__ __forceinline__ void put(long q, T v) const {
        pnew[q] = v;
        x[q] = (init ? X(0) : x[q]) + X(*alpha) * X(pold[q]);
    }
    template <int V>
    __device__ __forceinline__ void putv(long q, const TV<T, V>& u) const {
        stv<T, V>(pnew + q, u);
        const X al = X(*alpha);
        const TV<T, V> pp = ldv<T, V>(pold + q);
        // x may hold wider elements than p: V of them in one or more vector accesses
        constexpr int VX = 16 / sizeof(X) < V ? 16 / sizeof(X) : V;
#pragma unroll
        for (int h = 0; h < V; h += VX) {
            TV<X, VX> xx = xload<VX>(q + h);
#pragma unroll
            for (int e = 0; e < VX; ++e) xx.v[e] += al * X(pp.v[h + e]);
            stv<X, VX>(x + q + h, xx);
        }
    }
    static constexpr int NA = 2;
    __device__ __forceinline__ const T* lds_arr(int a) const { return a ? pold : z; }
    __device__ __forceinline__ T lds_scalar() const { return T(*beta); }
    __device__ __forceinline__ T lds_comb(T v0, T v1, T bt) const { return v0 + bt * v1; }
    // pp: p_old at the own points, as staged
    template <int V>
    __device__ __forceinline__ void lds_put(long q, const TV<T, V>& u, const TV<T, V>& pp) const {
        stv<T, V>(pnew + q, u);
        const X al = X(*alpha);
        constexpr int VX = 16 / sizeof(X) < V ? 16 / sizeof(X) : V;
#pragma unroll
        for (int h = 0; h < V; h += VX) {
            TV<X, VX> xx = xload<VX>(q + h);
#pragma unroll
            for (int e = 0; e < VX; ++e) xx.v[e] += al * X(pp.v[h + e]);
            stv<X, VX>(x + q + h, xx);
        }
    }
};

// u = r - alpha w, stored to rn: the CG residual update fused into the
// first kernel that reads the residual (the scale of a zero-guess smoothing
// step is applied by OutScaleU, so the stored value is the operand itself).
template <typename T>
struct SrcUpd {
    const T* __restrict__ r;
    const T* __restrict__ w;
    const double* __restrict__ alpha;  // device scalar
    T* __restrict__ rn;
    __device__ __forceinline__ T load(long q) const { return r[q] - T(*alpha) * w[q]; }
    template <int V>
    __device__ __forceinline__ TV<T, V> loadv(long q) const {
        const T al = T(*alpha);
        TV<T, V> rr = ldv<T, V>(r + q);
        const TV<T, V> ww = ldv<T, V>(w + q);
#pragma unroll
        for (int e = 0; e < V; ++e) rr.v[e] -= al * ww.v[e];
        return rr;
    }
    __device__ __forceinline__ void put(long q, T v) const { rn[q] = v; }
    template <int V>
    __device__ __forceinline__ void putv(long q, const TV<T, V>& u) const { stv<T, V>(rn + q, u); }
    static constexpr int NA = 2;
    __device__ __forceinline__ const T* lds_arr(int a) const { return a ? w : r; }
    __device__ __forceinline__ T lds_scalar() const { return T(*alpha); }
    __device__ __forceinline__ T lds_comb(T v0, T v1, T al) const { return v0 - al * v1; }
    template <int V>
    __device__ __forceinline__ void lds_put(long q, const TV<T, V>& u, const TV<T, V>&) const { stv<T, V>(rn + q, u); }
};

// u = x + alpha p (X wide, P narrow), stored to xnew: a pending iterate
// update folded into an FP64 residual pass.
template <typename X, typename P>
struct SrcAxpyPut {
    const X* __restrict__ x;
    const P* __restrict__ p;
    const double* __restrict__ alpha;  // device scalar
    X* __restrict__ xnew;
    __device__ __forceinline__ X load(long q) const { return x[q] + X(*alpha) * X(p[q]); }
    template <int V>
    __device__ __forceinline__ TV<X, V> loadv(long q) const {
        const X al = X(*alpha);
        TV<X, V> xx = ldv<X, V>(x + q);
        const TV<P, V> pp = ldv<P, V>(p + q);
#pragma unroll
        for (int e = 0; e < V; ++e) xx.v[e] += al * X(pp.v[e]);
        return xx;
    }
    __device__ __forceinline__ void put(long q, X v) const { xnew[q] = v; }
    template <int V>
    __device__ __forceinline__ void putv(long q, const TV<X, V>& u) const { stv<X, V>(xnew + q, u); }
};

// u = x + e + alpha p (X wide; e, p narrow), stored to xnew: an FP32
// correction and a pending iterate update folded into an FP64 residual pass.
template <typename X, typename P>
struct SrcAxpySumPut {
    const X* __restrict__ x;
    const P* __restrict__ e;
    const P* __restrict__ p;
    const double* __restrict__ alpha;  // device scalar
    X* __restrict__ xnew;
    __device__ __forceinline__ X load(long q) const { return x[q] + X(e[q]) + X(*alpha) * X(p[q]); }
    template <int V>
    __device__ __forceinline__ TV<X, V> loadv(long q) const {
        const X al = X(*alpha);
        TV<X, V> xx = ldv<X, V>(x + q);
        const TV<P, V> ee = ldv<P, V>(e + q);
        const TV<P, V> pp = ldv<P, V>(p + q);
#pragma unroll
        for (int k = 0; k < V; ++k) xx.v[k] = xx.v[k] + X(ee.v[k]) + al * X(pp.v[k]);
        return xx;
    }
    __device__ __forceinline__ void put(long q, X v) const { xnew[q] = v; }
    template <int V>
    __device__ __forceinline__ void putv(long q, const TV<X, V>& u) const { stv<X, V>(xnew + q, u); }
};

// the CG directions p_j (FP32 or FP64) and step lengths alpha_j (device
// scalars) of the iterations whose iterate update is deferred, in
// iteration order
constexpr int kFoldMax = 16;
template <typename P>
struct FoldListT {
    const P* p[kFoldMax];
    const double* a[kFoldMax];
    int m;
};
typedef FoldListT<float> FoldList;

// ------------------------------------------------------ output policies
//
// An output consumes u, A u and the Jacobi factor at a point (point) or at
// V consecutive points (pointv) and accumulates up to NDOT dot products.
// An output that reads NPRE per-point arrays besides its result lets the
// marching pass load them ahead: pre / pre1 load them at q (V points or
// one), pointvp / pointp consume them.  Those arrays must not be written by
// the pass's source.

template <typename T>
struct OutApplyDot {                  // w = A u ; acc0 += u.w
    T* __restrict__ w;
    static constexpr int NDOT = 1, NPRE = 0;
    __device__ __forceinline__ void point(long q, T u, T Au, T, double& a0, double&) const {
        w[q] = Au;
        a0 += double(u) * double(Au);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>&,
                                           double& a0, double&) const {
        stv<T, V>(w + q, Au);
#pragma unroll
        for (int e = 0; e < V; ++e) a0 += double(u.v[e]) * double(Au.v[e]);
    }
};

template <typename T>
struct OutApply {                     // w = A u
    T* __restrict__ w;
    static constexpr int NDOT = 0, NPRE = 0;
    __device__ __forceinline__ void point(long q, T, T Au, T, double&, double&) const { w[q] = Au; }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>&, const TV<T, V>& Au, const TV<T, V>&, double&,
                                           double&) const {
        stv<T, V>(w + q, Au);
    }
};

template <typename T, int NORM>
struct OutResid {                     // r = b - A u ; acc0 += r.r (NORM)
    const T* __restrict__ b;
    T* __restrict__ r;
    static constexpr int NDOT = NORM, NPRE = 1;
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const { p[0] = ldv<T, V>(b + q); }
    __device__ __forceinline__ void pre1(long q, T* p) const { p[0] = b[q]; }
    __device__ __forceinline__ void pointp(long q, T, T Au, T, const T* p, double& a0, double&) const {
        const T v = p[0] - Au;
        r[q] = v;
        if (NORM) a0 += double(v) * double(v);
    }
    template <int V>
    __device__ __forceinline__ void pointvp(long q, const TV<T, V>&, const TV<T, V>& Au, const TV<T, V>&,
                                            const TV<T, V>* p, double& a0, double&) const {
        TV<T, V> v;
#pragma unroll
        for (int e = 0; e < V; ++e) v.v[e] = p[0].v[e] - Au.v[e];
        stv<T, V>(r + q, v);
        if (NORM)
#pragma unroll
            for (int e = 0; e < V; ++e) a0 += double(v.v[e]) * double(v.v[e]);
    }
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        T p[1];
        pre1(q, p);
        pointp(q, u, Au, idg, p, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        TV<T, V> p[1];
        pre<V>(q, p);
        pointvp<V>(q, u, Au, idg, p, a0, a1);
    }
};

template <typename T>
struct OutResidNorms {                // acc0 += (b - A u).(b - A u), acc1 += b.b (nothing stored)
    const T* __restrict__ b;
    static constexpr int NDOT = 2, NPRE = 1;
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const { p[0] = ldv<T, V>(b + q); }
    __device__ __forceinline__ void pre1(long q, T* p) const { p[0] = b[q]; }
    __device__ __forceinline__ void pointp(long, T, T Au, T, const T* p, double& a0, double& a1) const {
        const T v = p[0] - Au;
        a0 += double(v) * double(v);
        a1 += double(p[0]) * double(p[0]);
    }
    template <int V>
    __device__ __forceinline__ void pointvp(long, const TV<T, V>&, const TV<T, V>& Au, const TV<T, V>&,
                                            const TV<T, V>* p, double& a0, double& a1) const {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const T v = p[0].v[e] - Au.v[e];
            a0 += double(v) * double(v);
            a1 += double(p[0].v[e]) * double(p[0].v[e]);
        }
    }
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        T p[1];
        pre1(q, p);
        pointp(q, u, Au, idg, p, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        TV<T, V> p[1];
        pre<V>(q, p);
        pointvp<V>(q, u, Au, idg, p, a0, a1);
    }
};

template <typename T, typename R>
struct OutResidCvt {                  // r = b - A u in T, stored as R ; acc0 += r.r
    const T* __restrict__ b;
    R* __restrict__ r;
    static constexpr int NDOT = 1, NPRE = 1;
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const { p[0] = ldv<T, V>(b + q); }
    __device__ __forceinline__ void pre1(long q, T* p) const { p[0] = b[q]; }
    __device__ __forceinline__ void pointp(long q, T, T Au, T, const T* p, double& a0, double&) const {
        const T v = p[0] - Au;
        r[q] = R(v);
        a0 += double(v) * double(v);
    }
    template <int V>
    __device__ __forceinline__ void pointvp(long q, const TV<T, V>&, const TV<T, V>& Au, const TV<T, V>&,
                                            const TV<T, V>* p, double& a0, double&) const {
        TV<R, V> o;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const T v = p[0].v[e] - Au.v[e];
            o.v[e] = R(v);
            a0 += double(v) * double(v);
        }
        stv<R, V>(r + q, o);
    }
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        T p[1];
        pre1(q, p);
        pointp(q, u, Au, idg, p, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        TV<T, V> p[1];
        pre<V>(q, p);
        pointvp<V>(q, u, Au, idg, p, a0, a1);
    }
};

// Zero-guess smoothing step c b on a level whose right-hand side b is the
// operand u itself (read by the source, or formed and stored as b by it):
// the inner output sees c u, c A u, and u as its b (its one per-point array).
template <typename T, class Out>
struct OutScaleU {
    T c;
    Out inner;
    static_assert(Out::NPRE == 1, "OutScaleU supplies the inner output's b");
    static constexpr int NDOT = Out::NDOT, NPRE = 0;
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        inner.pointp(q, c * u, c * Au, idg, &u, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        TV<T, V> cu, cAu;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            cu.v[e] = c * u.v[e];
            cAu.v[e] = c * Au.v[e];
        }
        inner.template pointvp<V>(q, cu, cAu, idg, &u, a0, a1);
    }
};

// Wraps an output without dot products and adds acc0 += u.u at every point
// (u: the residual formed and stored by SrcUpd).
template <typename T, class Out>
struct OutRR {
    Out inner;
    static_assert(Out::NDOT == 0 && Out::NPRE == 0, "OutRR wraps an output without dots or operands");
    static constexpr int NDOT = 1, NPRE = 0;
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        inner.point(q, u, Au, idg, a0, a1);
        a0 += double(u) * double(u);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        inner.template pointv<V>(q, u, Au, idg, a0, a1);
#pragma unroll
        for (int e = 0; e < V; ++e) a0 += double(u.v[e]) * double(u.v[e]);
    }
};

// Stores the operand u at every point (the prolonged iterate t of a fused
// prolongation + smoothing pass, kept for the next Chebyshev step), then the
// inner output.
template <typename T, class Out>
struct OutKeepU {
    T* __restrict__ keep;
    Out inner;
    static constexpr int NDOT = Out::NDOT, NPRE = Out::NPRE;
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const { inner.template pre<V>(q, p); }
    __device__ __forceinline__ void pre1(long q, T* p) const { inner.pre1(q, p); }
    __device__ __forceinline__ void pointp(long q, T u, T Au, T idg, const T* p, double& a0, double& a1) const {
        keep[q] = u;
        inner.pointp(q, u, Au, idg, p, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointvp(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                            const TV<T, V>* p, double& a0, double& a1) const {
        stv<T, V>(keep + q, u);
        inner.template pointvp<V>(q, u, Au, idg, p, a0, a1);
    }
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        keep[q] = u;
        inner.point(q, u, Au, idg, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        stv<T, V>(keep + q, u);
        inner.template pointv<V>(q, u, Au, idg, a0, a1);
    }
};

// y = c0 prev + c1 u + c2 D^-1 (b - A u); DOTS: acc0 += y.b, acc1 += y.y
template <typename T, bool PREV, int DOTS>
struct OutCheb {
    const T* __restrict__ b;
    const T* __restrict__ prev;
    T* __restrict__ y;
    T c0, c1, c2;
    static constexpr int NDOT = DOTS, NPRE = PREV ? 2 : 1;   // b, prev
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const {
        p[0] = ldv<T, V>(b + q);
        if (PREV) p[NPRE - 1] = ldv<T, V>(prev + q);
    }
    __device__ __forceinline__ void pre1(long q, T* p) const {
        p[0] = b[q];
        if (PREV) p[NPRE - 1] = prev[q];
    }
    __device__ __forceinline__ void pointp(long q, T u, T Au, T idg, const T* p, double& a0, double& a1) const {
        T v = c1 * u + c2 * (idg * (p[0] - Au));
        if (PREV) v += c0 * p[NPRE - 1];
        y[q] = v;
        if (DOTS) {
            a0 += double(v) * double(p[0]);
            a1 += double(v) * double(v);
        }
    }
    template <int V>
    __device__ __forceinline__ void pointvp(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                            const TV<T, V>* p, double& a0, double& a1) const {
        TV<T, V> v;
#pragma unroll
        for (int e = 0; e < V; ++e) v.v[e] = c1 * u.v[e] + c2 * (idg.v[e] * (p[0].v[e] - Au.v[e]));
        if (PREV)
#pragma unroll
            for (int e = 0; e < V; ++e) v.v[e] += c0 * p[NPRE - 1].v[e];
        stv<T, V>(y + q, v);
        if (DOTS)
#pragma unroll
            for (int e = 0; e < V; ++e) {
                a0 += double(v.v[e]) * double(p[0].v[e]);
                a1 += double(v.v[e]) * double(v.v[e]);
            }
    }
    __device__ __forceinline__ void point(long q, T u, T Au, T idg, double& a0, double& a1) const {
        T p[NPRE];
        pre1(q, p);
        pointp(q, u, Au, idg, p, a0, a1);
    }
    template <int V>
    __device__ __forceinline__ void pointv(long q, const TV<T, V>& u, const TV<T, V>& Au, const TV<T, V>& idg,
                                           double& a0, double& a1) const {
        TV<T, V> p[NPRE];
        pre<V>(q, p);
        pointvp<V>(q, u, Au, idg, p, a0, a1);
    }
};

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// fixed-order block sum of up to two values; thread 0 writes part[2*bid+q]
// (blocks of at most NW waves)
template <int NDOT, int NW = NT / 64>
__device__ __forceinline__ void block_partials(double a0, double a1, double* __restrict__ part,
                                               long bid) {
    if (NDOT == 0) return;
    __shared__ double sm[2][NW];
    const int lin = threadIdx.x + blockDim.x * threadIdx.y;
    const int wv = lin >> 6, ln = lin & 63;
    const int nw = (blockDim.x * blockDim.y) >> 6;
    a0 = wave_sum(a0);
    if (NDOT > 1) a1 = wave_sum(a1);
    if (ln == 0) { sm[0][wv] = a0; sm[1][wv] = a1; }
    __syncthreads();
    if (lin == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int q = 0; q < nw; ++q) { s0 += sm[0][q]; s1 += sm[1][q]; }
        part[2 * bid] = s0;
        part[2 * bid + 1] = s1;
    }
}

// ------------------------------------------------------ block placement
//
// The dispatcher places workgroup b (x fastest, then y, z) on XCD b mod 8;
// each XCD has its own L2.  Neighbouring tiles of a marching pass share halo
// rows and planes, so the launch's tiles are renumbered: XCD x takes the
// contiguous tile range [x q + min(x, r), ...) of q or q + 1 tiles
// (q = n / 8, r = n mod 8), visited in dispatch order.  The map is a
// bijection of [0, n) for every n.
struct Tile {
    int bx, by, bz;
    long id;            // renumbered linear index (partial-sum slot)
};

__device__ __forceinline__ Tile xcd_tile() {
    const long n = (long)gridDim.x * gridDim.y * gridDim.z;
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    const long q = n >> 3, r = n & 7, x = b & 7;
    const long t = x * q + (x < r ? x : r) + (b >> 3);
    Tile o;
    o.id = t;
    o.bx = (int)(t % gridDim.x);
    const long u = t / gridDim.x;
    o.by = (int)(u % gridDim.y);
    o.bz = (int)(u / gridDim.y);
    return o;
}

// ------------------------------------------------------------- stencil

// Marching stencil pass over a 2D or 3D grid: for every vertex, u from Src,
// A u and the Jacobi factor from Op, result written by Out.  A lane owns V
// consecutive x points i0 .. i0+V-1, i0 = 1 + V (64 bx + lane), moved with
// one 16-byte vector access (V = 2 in FP64, 4 in FP32); the west value of
// i0 and the east value of i0+V-1 come from the neighbouring lanes (the
// wave's edge lanes load them).  The lane at i0 = 1 also owns the boundary
// point i = 0; a lane at the end of a row whose points do not fill V uses
// element accesses.  The march runs along the slowest dimension (z in 3D,
// y in 2D) keeping the planes m-1, m, m+1 in registers; 3D rows j+-1 are
// read through the cache.  Every lane of a wave runs the same loop (loads
// are predicated) so the cross-lane exchanges see the whole wave.
// Latency hiding: the centre row of plane m+PD is loaded while plane m is
// computed (a ring of PD plane records: the row, the edge lanes' extra x
// value with PF bit 1, the output's per-point operands); PF bit 0 loads rows
// j+-1 of plane m+1 one plane early.
// Grid:
//   3D: (ceil((mx-1)/(64V)), ceil(my/TY), ceil(mz/kch)), block (64,TY)
//   2D: (ceil((mx-1)/(64V)), ceil(ceil(my/kch)/TY), 1),  block (64,TY)
template <typename T, int DIM, int V, class Op, class Src, class Out, int PD = 1, int PF = 3>
__device__ __forceinline__ void stencil_vec(int bx, int by, int bz, int mx, int my, int mz, int kch,
                                            const Op& op, const Src& src, const Out& out, double& a0,
                                            double& a1) {
    static_assert(PD >= 1 && PD <= 4, "prefetch distance");
    constexpr bool PSN = (PF & 1) != 0 && DIM == 3, PEX = (PF & 2) != 0;
    constexpr int NP = Out::NPRE, NPA = NP > 0 ? NP : 1;
    const int lane = threadIdx.x;
    const int i0 = 1 + V * (bx * TX + lane);
    const T z0 = T(0);
    const long sy = mx;
    int j, m0, m1, M;
    long sm, base;
    if (DIM == 3) {
        j = by * TY + threadIdx.y;
        M = mz;
        sm = (long)mx * my;
        m0 = bz * kch;
        m1 = min(m0 + kch, mz);
        base = (long)j * sy;
        if (j >= my) return;           // uniform per wave
    } else {
        j = 0;
        M = my;
        sm = sy;
        m0 = (by * TY + threadIdx.y) * kch;
        m1 = min(m0 + kch, my);
        base = 0;
    }
    if (m0 >= m1) return;              // uniform per wave
    const bool act = i0 < mx;
    const int ne = act ? min(V, mx - i0) : 0;     // points of this lane
    const bool full = ne == V;
    const bool jb = DIM == 3 && (j == 0 || j == my - 1);
    const bool ms = DIM == 3 && j - 1 > 0, mn = DIM == 3 && j + 1 < my - 1;
    const bool wl = lane == 0 && i0 - 1 > 0, el = lane == TX - 1 && i0 + V < mx - 1;
    TV<T, V> zv;
#pragma unroll
    for (int e = 0; e < V; ++e) zv.v[e] = z0;
    // (active lanes only: ne >= 1, so element 0 of a partial lane is in range)
    auto LV = [&](long a) -> TV<T, V> {
        if (full) return src.template loadv<V>(a);
        TV<T, V> r;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const T v = src.load(a + (e < ne ? e : 0));
            r.v[e] = e < ne ? v : z0;
        }
        return r;
    };
    // a plane's values on the centre row: the operand (u), the edge lanes'
    // extra value (x: lane 0 i0-1, last lane i0+V when interior; PEX), the
    // output's per-point operands (o: full lanes; ox: at i0-1 = 0 for the
    // lane with i0 = 1).  c*: plane m; r*[d]: plane m+1+d.
    TV<T, V> cu = zv, ru[PD], co[NPA], ro[PD][NPA];
    T cx = z0, rx[PD], cox[NPA], rox[PD][NPA];
#pragma unroll
    for (int k = 0; k < NPA; ++k) {
        co[k] = zv;
        cox[k] = z0;
    }
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        ru[d] = zv;
        rx[d] = z0;
#pragma unroll
        for (int k = 0; k < NPA; ++k) {
            ro[d][k] = zv;
            rox[d][k] = z0;
        }
    }
    // load the values of the plane at a (own: the plane is computed here)
    auto LR = [&](long a, bool own, TV<T, V>& u, T& x, TV<T, V>* o, T* ox) {
        u = LV(a);
        if (PEX) x = lane == 0 ? src.load(a - 1) : (el ? src.load(a + V) : z0);
        if constexpr (NP > 0) {
            if (own && full) out.template pre<V>(a, o);
            if (own && i0 == 1) out.pre1(a - 1, ox);
        }
    };
    // rows j-1, j+1 of plane m are read by interior rows of interior planes
    auto snok = [&](int m) -> bool { return PSN && act && !jb && m > 0 && m < M - 1; };
    const int mlast = min(m1, M - 1);  // last plane whose centre row is read
    long q = base + (long)m0 * sm + i0;
    TV<T, V> ud = zv;
    TV<T, V> sc = zv, nc = zv;
    if (act) {
        if (m0 > 0) ud = LV(q - sm);
        LR(q, true, cu, cx, co, cox);
#pragma unroll
        for (int d = 0; d + 1 < PD; ++d)
            if (m0 + 1 + d <= mlast) LR(q + (1 + d) * sm, m0 + 1 + d < m1, ru[d], rx[d], ro[d], rox[d]);
        if (snok(m0)) {
            if (ms) sc = LV(q - sy);
            if (mn) nc = LV(q + sy);
        }
    }
    for (int m = m0; m < m1; ++m, q += sm) {
        if (act && m + PD <= mlast)
            LR(q + PD * sm, m + PD < m1, ru[PD - 1], rx[PD - 1], ro[PD - 1], rox[PD - 1]);
        TV<T, V> sn = zv, nn = zv;
        if (m + 1 < m1 && snok(m + 1)) {
            if (ms) sn = LV(q + sm - sy);
            if (mn) nn = LV(q + sm + sy);
        }
        const TV<T, V> uc = cu;
        const TV<T, V> uu = ru[0];     // plane m+1 (read only when m+1 < M)
        T wv = __shfl_up(uc.v[V - 1], 1, TX);
        T ev = __shfl_down(uc.v[0], 1, TX);
        if (act) {
            if (PEX) {
                if (wl) wv = cx;
                if (el) ev = cx;
            } else {
                if (wl) wv = src.load(q - 1);
                if (el) ev = src.load(q + V);
            }
            const int jj = DIM == 3 ? j : m, kk = DIM == 3 ? m : 0;
            if (full) {
                src.template putv<V>(q, uc);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (e < ne) src.put(q + e, uc.v[e]);
            }
            TV<T, V> Au = zv, idg = zv;
            if (jb || m == 0 || m == M - 1) {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (e < ne) op.bnd(i0 + e, jj, kk, uc.v[e], Au.v[e], idg.v[e]);
            } else {
                TV<T, V> sv = zv, nv = zv, dn = zv, up = zv;
                if (DIM == 3) {
                    if (PSN) {
                        sv = sc;
                        nv = nc;
                    } else {
                        if (ms) sv = LV(q - sy);
                        if (mn) nv = LV(q + sy);
                    }
                    if (m - 1 > 0) dn = ud;
                    if (m + 1 < M - 1) up = uu;
                } else {
                    if (m - 1 > 0) sv = ud;
                    if (m + 1 < M - 1) nv = uu;
                }
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const int i = i0 + e;
                    if (e < ne) {
                        if (i == mx - 1) {
                            op.bnd(i, jj, kk, uc.v[e], Au.v[e], idg.v[e]);
                        } else {
                            const T w = e == 0 ? (i - 1 > 0 ? wv : z0) : uc.v[e > 0 ? e - 1 : 0];
                            const T east = i + 1 < mx - 1 ? (e == V - 1 ? ev : uc.v[e < V - 1 ? e + 1 : 0]) : z0;
                            op.interior(i, jj, kk, uc.v[e], w, east, sv.v[e], nv.v[e], dn.v[e], up.v[e], Au.v[e],
                                        idg.v[e]);
                        }
                    }
                }
            }
            if (full) {
                if constexpr (NP > 0) out.template pointvp<V>(q, uc, Au, idg, co, a0, a1);
                else out.template pointv<V>(q, uc, Au, idg, a0, a1);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (e < ne) out.point(q + e, uc.v[e], Au.v[e], idg.v[e], a0, a1);
            }
            if (i0 == 1) {
                const T u0 = PEX ? cx : src.load(q - 1);
                T A0, i00;
                src.put(q - 1, u0);
                op.bnd(0, jj, kk, u0, A0, i00);
                if constexpr (NP > 0) out.pointp(q - 1, u0, A0, i00, cox, a0, a1);
                else out.point(q - 1, u0, A0, i00, a0, a1);
            }
        }
        ud = cu;
        cu = ru[0];
        cx = rx[0];
#pragma unroll
        for (int k = 0; k < NPA; ++k) {
            co[k] = ro[0][k];
            cox[k] = rox[0][k];
        }
#pragma unroll
        for (int d = 0; d + 1 < PD; ++d) {
            ru[d] = ru[d + 1];
            rx[d] = rx[d + 1];
#pragma unroll
            for (int k = 0; k < NPA; ++k) {
                ro[d][k] = ro[d + 1][k];
                rox[d][k] = rox[d + 1][k];
            }
        }
        sc = sn;
        nc = nn;
    }
}

// Generic stencil pass (see stencil_vec); 1D: thread per point from i = 1,
// grid (ceil((mx-1)/64), 1, 1), block (64,1).
template <typename T, int DIM, class Op, class Src, class Out>
__device__ __forceinline__ void stencil_body(int bx, int by, int bz, int mx, int my, int mz, int kch,
                                             const Op& op, const Src& src, const Out& out,
                                             double& a0, double& a1) {
    if (DIM != 1) {
        stencil_vec<T, DIM, VecOf<T>::V, Op, Src, Out>(bx, by, bz, mx, my, mz, kch, op, src, out, a0, a1);
        return;
    }
    const int i = 1 + bx * TX + threadIdx.x;
    const T z0 = T(0);
    if (i < mx) {
        const long q = i;
        const T uc = src.load(q);
        src.put(q, uc);
        T Au, idg;
        if (i == mx - 1) {
            op.bnd(i, 0, 0, uc, Au, idg);
        } else {
            const T w = (i - 1 > 0) ? src.load(q - 1) : z0;
            const T e = (i + 1 < mx - 1) ? src.load(q + 1) : z0;
            op.interior(i, 0, 0, uc, w, e, z0, z0, z0, z0, Au, idg);
        }
        out.point(q, uc, Au, idg, a0, a1);
        if (i == 1) {
            const T u0 = src.load(0);
            src.put(0, u0);
            op.bnd(0, 0, 0, u0, Au, idg);
            out.point(0, u0, Au, idg, a0, a1);
        }
    }
}

template <typename T, int DIM, class Op, class Src, class Out>
__global__ void __launch_bounds__(NT)
k_stencil(int mx, int my, int mz, int kch, Op op, Src src, Out out, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    const Tile tl = xcd_tile();
    stencil_body<T, DIM, Op, Src, Out>(tl.bx, tl.by, tl.bz, mx, my, mz, kch, op, src, out, a0, a1);
    block_partials<Out::NDOT>(a0, a1, part, tl.id);
}

// ------------------------------------------------ LDS-staged stencil pass
//
// The 3D marching pass of stencil_vec with the source's arrays staged in
// LDS.  A block of TYD waves covers the rows j0 .. j0+TYD-1 (j0 = TYD by) of
// W = 64 V points (i = is .. is+W-1, is = 1 + W bx) and marches the planes
// k0 .. k1-1.  Every plane's rows j0-1 .. j0+TYD of each array the source
// reads at stencil neighbours go into a ring of S = P + 2 LDS slots by
// LDS-DMA loads issued P planes ahead of use: one global_load_lds_dwordx4
// lands a row's W values (64 lanes x 16 bytes), one 4-byte load per lane
// the two x-halo values (is-1 and is+W) a dword at a time.  A wave waits for
// its own loads of the planes it needs (s_waitcnt vmcnt, counting only
// those loads, so the wait also covers any later memory operation) and the
// block's barrier makes every wave's rows visible; a slot is refilled only
// after the barrier that follows its last read.  The operand of the own row
// at planes k-1, k is kept in registers (at a chunk's first plane, plane
// k0-1 is read from memory by the source), the plane k+1 and rows j+-1 come
// from LDS, so each staged value is read from memory once per block where
// stencil_vec reads rows j+-1 again through the cache.  The output's
// per-point operands are plain loads one plane ahead.  The arithmetic per
// point is stencil_vec's (same operand formula, same operation order).
// Requires 3D, (mx - 1) % W == 0 and 4- or 8-byte elements.
// Grid (ceil((mx-1)/W), ceil(my/TYD), ceil(mz/kch)), block (64, TYD).

// s_waitcnt vmcnt(n) alone (the other counters not waited for)
#define P4G_VMW(N)                                          \
    case N:                                                 \
        asm volatile("s_waitcnt vmcnt(" #N ")" ::: "memory"); \
        break;
__device__ __forceinline__ void wait_vmcnt(int n) {
    switch (n) {
        P4G_VMW(0) P4G_VMW(1) P4G_VMW(2) P4G_VMW(3) P4G_VMW(4) P4G_VMW(5) P4G_VMW(6) P4G_VMW(7) P4G_VMW(8)
        P4G_VMW(9) P4G_VMW(10) P4G_VMW(11) P4G_VMW(12) P4G_VMW(13) P4G_VMW(14) P4G_VMW(15) P4G_VMW(16)
        default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
}
#undef P4G_VMW
// a workgroup barrier that orders the surrounding memory operations in the
// compiler without waiting for outstanding loads (the waits are explicit)
__device__ __forceinline__ void lds_barrier() { asm volatile("s_barrier" ::: "memory"); }

typedef __attribute__((address_space(3))) void lds_void_t;

template <typename T, class Op, class Src, class Out, int TYD, int P>
__device__ __forceinline__ void stencil_lds(int bx, int by, int bz, int mx, int my, int mz, int kch, const Op& op,
                                            const Src& src, const Out& out, double& a0, double& a1) {
    constexpr int V = VecOf<T>::V, W = TX * V, S = P + 2, NA = Src::NA;
    constexpr int DPE = sizeof(T) / 4;               // dwords per element
    constexpr int NP = Out::NPRE, NPA = NP > 0 ? NP : 1;
    static_assert(V * sizeof(T) == 16 && (DPE == 1 || DPE == 2), "16-byte lanes of 4- or 8-byte elements");
    __shared__ __attribute__((aligned(16))) T ring[NA][S][TYD + 2][W];
    __shared__ __attribute__((aligned(16))) T halo[NA][S][TYD + 2][2];
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int is = 1 + W * bx;
    const int j0 = TYD * by;
    const int k0 = bz * kch, k1 = min(k0 + kch, mz);
    if (k0 >= k1) return;                            // uniform per block
    const int klast = min(k1, mz - 1);               // last plane read
    // the rows this wave stages: r = w and r = w + TYD (rows j0 - 1 + r)
    const int ra = w, rb = w + TYD;
    const bool oka = j0 - 1 + ra >= 0 && j0 - 1 + ra < my;
    const bool okb = rb < TYD + 2 && j0 - 1 + rb < my;
    const int nload = 2 * NA * ((int)oka + (int)okb);   // DMA instructions per plane
    const bool has_e = is + W < mx;
    const long sm = (long)mx * my;
    auto stage_row = [&](int kk, int r) {
        const int slot = (kk - k0 + S) % S;
        const long rowq = (long)kk * sm + (long)(j0 - 1 + r) * mx;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const T* g = src.lds_arr(a) + rowq;
            __builtin_amdgcn_global_load_lds((const void*)(g + is + V * lane), (lds_void_t*)&ring[a][slot][r][0], 16,
                                             0, 0);
            if (lane < 2 * DPE) {                   // W dwords on lanes [0, DPE), E dwords on [DPE, 2 DPE)
                const bool east = lane >= DPE;
                if (!east || has_e)
                    __builtin_amdgcn_global_load_lds(
                        (const void*)((const char*)(g + (east ? is + W : is - 1)) + 4 * (lane % DPE)),
                        (lds_void_t*)&halo[a][slot][r][0], 4, 0, 0);
            }
        }
    };
    auto stage = [&](int kk) {
        if (kk < 0 || kk > klast) return;
        if (oka) stage_row(kk, ra);
        if (okb) stage_row(kk, rb);
    };
    for (int kk = k0; kk <= k0 + P; ++kk) stage(kk);

    const int j = j0 + w;
    const bool rowok = j < my;                       // uniform per wave
    const bool jb = j == 0 || j == my - 1;
    const bool ms = j - 1 > 0, mn = j + 1 < my - 1;
    const int r = w + 1;
    const int i0 = is + V * lane;
    const T z0 = T(0);
    const T sc_ = src.lds_scalar();
    // the operand at LDS row rr of slot sl: V values of the lane, one element
    auto opv = [&](int sl, int rr) -> TV<T, V> {
        TV<T, V> v = ldv<T, V>(&ring[0][sl][rr][V * lane]);
        if constexpr (NA == 2) {
            const TV<T, V> v1 = ldv<T, V>(&ring[1][sl][rr][V * lane]);
#pragma unroll
            for (int e = 0; e < V; ++e) v.v[e] = src.lds_comb(v.v[e], v1.v[e], sc_);
        }
        return v;
    };
    auto op1 = [&](int sl, int rr, int idx) -> T {
        T v = ring[0][sl][rr][idx];
        if constexpr (NA == 2) v = src.lds_comb(v, ring[1][sl][rr][idx], sc_);
        return v;
    };
    auto oph = [&](int sl, int rr, int e) -> T {
        T v = halo[0][sl][rr][e];
        if constexpr (NA == 2) v = src.lds_comb(v, halo[1][sl][rr][e], sc_);
        return v;
    };
    TV<T, V> zv;
#pragma unroll
    for (int e = 0; e < V; ++e) zv.v[e] = z0;
    TV<T, V> co[NPA], ro[NPA];
    T cox[NPA], rox[NPA];
#pragma unroll
    for (int q = 0; q < NPA; ++q) {
        co[q] = ro[q] = zv;
        cox[q] = rox[q] = z0;
    }
    auto pre_own = [&](long qq, TV<T, V>* o, T* ox) {
        if constexpr (NP > 0) {
            out.template pre<V>(qq, o);
            if (i0 == 1) out.pre1(qq - 1, ox);
        }
    };
    long q = (long)k0 * sm + (long)j * mx + i0;
    if (rowok) pre_own(q, co, cox);
    TV<T, V> ud = zv, uc = zv;       // the own row's operand at planes k-1, k
    if (rowok && k0 > 0) ud = src.template loadv<V>(q - sm);
    for (int k = k0; k < k1; ++k, q += sm) {
        // planes up to k+1 staged: at most the planes k+2 .. min(k+P, klast) outstanding
        wait_vmcnt(max(0, min(k + P, klast) - (k + 1)) * nload);
        lds_barrier();
        stage(k + 1 + P);
        if (!rowok) continue;                        // uniform per wave (no barrier skipped)
        if (k + 1 < k1) pre_own(q + sm, ro, rox);
        const int sc = (k - k0) % S, su = (k - k0 + 1) % S;
        if (k == k0) uc = opv(sc, r);
        const TV<T, V> uu = k + 1 < mz ? opv(su, r) : zv;   // plane k+1 (read only when k+1 < mz)
        TV<T, V> Au = zv, idg = zv;
        if (jb || k == 0 || k == mz - 1) {
#pragma unroll
            for (int e = 0; e < V; ++e) op.bnd(i0 + e, j, k, uc.v[e], Au.v[e], idg.v[e]);
        } else {
            const TV<T, V> sv = ms ? opv(sc, r - 1) : zv, nv = mn ? opv(sc, r + 1) : zv;
            const TV<T, V> dn = k - 1 > 0 ? ud : zv, up = k + 1 < mz - 1 ? uu : zv;
            const T wv = lane > 0 ? op1(sc, r, V * lane - 1) : oph(sc, r, 0);
            const T ev = lane < TX - 1 ? op1(sc, r, V * lane + V) : oph(sc, r, 1);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int i = i0 + e;
                if (i == mx - 1) {
                    op.bnd(i, j, k, uc.v[e], Au.v[e], idg.v[e]);
                } else {
                    const T wn = e == 0 ? (i - 1 > 0 ? wv : z0) : uc.v[e > 0 ? e - 1 : 0];
                    const T east = i + 1 < mx - 1 ? (e == V - 1 ? ev : uc.v[e < V - 1 ? e + 1 : 0]) : z0;
                    op.interior(i, j, k, uc.v[e], wn, east, sv.v[e], nv.v[e], dn.v[e], up.v[e], Au.v[e], idg.v[e]);
                }
            }
        }
        if constexpr (NA == 2) src.template lds_put<V>(q, uc, ldv<T, V>(&ring[1][sc][r][V * lane]));
        else src.template lds_put<V>(q, uc, uc);
        if constexpr (NP > 0) out.template pointvp<V>(q, uc, Au, idg, co, a0, a1);
        else out.template pointv<V>(q, uc, Au, idg, a0, a1);
        if (i0 == 1) {                               // the boundary point i = 0
            const T u0 = oph(sc, r, 0);
            T A0, i00;
            src.put(q - 1, u0);
            op.bnd(0, j, k, u0, A0, i00);
            if constexpr (NP > 0) out.pointp(q - 1, u0, A0, i00, cox, a0, a1);
            else out.point(q - 1, u0, A0, i00, a0, a1);
        }
        ud = uc;
        uc = uu;
#pragma unroll
        for (int qq = 0; qq < NPA; ++qq) {
            co[qq] = ro[qq];
            cox[qq] = rox[qq];
        }
    }
}

// sources with LDS staging (those that define NA)
template <class S, class = void>
struct lds_stageable {
    static constexpr bool value = false;
};
template <class S>
struct lds_stageable<S, decltype((void)S::NA)> {
    static constexpr bool value = S::NA >= 1;
};

template <typename T, class Op, class Src, class Out, int TYD, int P>
__global__ void __launch_bounds__(TX * TYD)
k_stencil_lds(int mx, int my, int mz, int kch, Op op, Src src, Out out, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    const Tile tl = xcd_tile();
    stencil_lds<T, Op, Src, Out, TYD, P>(tl.bx, tl.by, tl.bz, mx, my, mz, kch, op, src, out, a0, a1);
    block_partials<Out::NDOT, TYD>(a0, a1, part, tl.id);
}

// ------------------------------------------- Galerkin coarse operators

// A level operator in 27-point stencil form: plane c = 9 (oz+1) + 3 (oy+1)
// + (ox+1) of g holds, for every point q, the coefficient of row q toward
// the point at offset (ox, oy, oz) (zero where that point is outside the
// grid).  Coarse operators A_c = P^T A P (-pc_mg_galerkin) are of this form.
template <typename T>
struct Op27 {
    const double* __restrict__ g;
    long n;
    int mx, my, mz;
    int jac;                     // Jacobi factor 1 / A_qq, else 1
};

// Generic pass on a 27-point level: a thread per point (grid-stride), the
// operand and its neighbours from Src, the row sum in column order.
// Launch: <<<blocks, 256>>>, one partial-sum slot per block.
template <typename T, class Src, class Out>
__global__ void __launch_bounds__(256)
k_stencil27(Op27<T> op, Src src, Out out, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    const long sxy = (long)op.mx * op.my;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < op.n; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % op.mx);
        const long r = q / op.mx;
        const int j = (int)(r % op.my), k = (int)(r / op.my);
        T Au = T(0);
        for (int oz = -1; oz <= 1; ++oz) {
            if (k + oz < 0 || k + oz >= op.mz) continue;
            for (int oy = -1; oy <= 1; ++oy) {
                if (j + oy < 0 || j + oy >= op.my) continue;
                for (int ox = -1; ox <= 1; ++ox) {
                    if (i + ox < 0 || i + ox >= op.mx) continue;
                    const int c = 9 * (oz + 1) + 3 * (oy + 1) + (ox + 1);
                    const T a = T(op.g[c * op.n + q]);
                    if (a != T(0)) Au += a * src.load(q + oz * sxy + oy * (long)op.mx + ox);
                }
            }
        }
        const T uc = src.load(q);
        src.put(q, uc);
        // a zero diagonal marks a null row (an active vertex of a reduced-
        // space VI step): its Jacobi factor is 0, so smoothing leaves it
        const T dq = T(op.g[13 * op.n + q]);
        const T idg = op.jac ? (dq != T(0) ? T(1) / dq : T(0)) : T(1);
        out.point(q, uc, Au, idg, a0, a1);
    }
    block_partials<Out::NDOT, 4>(a0, a1, part, blockIdx.x);
}

// The fine operator of a Galerkin product: the rediscretized constant
// 5/7-point operator (g == nullptr: Dirichlet rows diagonal-only, interior
// rows coupled to interior neighbours) or a 27-point level.
struct GalFine {
    int dim, mx, my, mz;
    double d, db, sx, sy, sz;
    const double* g;
    long n;
};

__device__ __forceinline__ bool gal_bnd(const GalFine& f, int i, int j, int k) {
    return i == 0 || i == f.mx - 1 || (f.dim >= 2 && (j == 0 || j == f.my - 1)) ||
           (f.dim == 3 && (k == 0 || k == f.mz - 1));
}

// coefficient of the fine operator in row (i, j, k) toward offset (ox, oy, oz)
__device__ __forceinline__ double gal_fine(const GalFine& f, int i, int j, int k, int ox, int oy, int oz) {
    if (f.g) return f.g[(long)(9 * (oz + 1) + 3 * (oy + 1) + (ox + 1)) * f.n + ((long)k * f.my + j) * f.mx + i];
    const bool bnd = gal_bnd(f, i, j, k);
    if (ox == 0 && oy == 0 && oz == 0) return bnd ? f.db : f.d;
    if (bnd || (ox != 0) + (oy != 0) + (oz != 0) != 1) return 0.0;
    if (gal_bnd(f, i + ox, j + oy, k + oz)) return 0.0;
    return ox ? -f.sx : (oy ? -f.sy : -f.sz);
}

// A_c = P^T A_f P in 27-point form, a thread per coarse point (grid-stride).
// P is the Q1 interpolation on all vertices (weights 1, 1/2 per axis, fine
// 2I <-> coarse I).  Coefficient (I, I+o) = sum over fine i near 2I and
// fine j = i + s near 2(I+o) of P(i,I) A_f(i,j) P(j,I+o).
__global__ void __launch_bounds__(256)
k_galerkin(GalFine f, int mxc, int myc, int mzc, double* __restrict__ gc) {
    const long nc = (long)mxc * myc * mzc;
    auto wt = [](int e) -> double { return e == 0 ? 1.0 : (e == 1 || e == -1 ? 0.5 : 0.0); };
    for (long q = blockIdx.x * 256L + threadIdx.x; q < nc; q += (long)gridDim.x * 256L) {
        const int I = (int)(q % mxc);
        const long r = q / mxc;
        const int J = (int)(r % myc), K = (int)(r / myc);
        for (int c = 0; c < 27; ++c) {
            const int oz = c / 9 - 1, oy = (c / 3) % 3 - 1, ox = c % 3 - 1;
            const int I2 = I + ox, J2 = J + oy, K2 = K + oz;
            double acc = 0.0;
            if (I2 >= 0 && I2 < mxc && J2 >= 0 && J2 < myc && K2 >= 0 && K2 < mzc) {
                for (int dz = -1; dz <= 1; ++dz) {
                    const int fk = 2 * K + dz;
                    if (fk < 0 || fk >= f.mz) continue;
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int fj = 2 * J + dy;
                        if (fj < 0 || fj >= f.my) continue;
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int fi = 2 * I + dx;
                            if (fi < 0 || fi >= f.mx) continue;
                            const double wi = wt(dx) * wt(dy) * wt(dz);
                            for (int sz = -1; sz <= 1; ++sz) {
                                const int gk = fk + sz, ez = gk - 2 * K2;
                                if (gk < 0 || gk >= f.mz || ez < -1 || ez > 1) continue;
                                for (int sy = -1; sy <= 1; ++sy) {
                                    const int gj = fj + sy, ey = gj - 2 * J2;
                                    if (gj < 0 || gj >= f.my || ey < -1 || ey > 1) continue;
                                    for (int sx = -1; sx <= 1; ++sx) {
                                        const int gi = fi + sx, ex = gi - 2 * I2;
                                        if (gi < 0 || gi >= f.mx || ex < -1 || ex > 1) continue;
                                        const double a = gal_fine(f, fi, fj, fk, sx, sy, sz);
                                        if (a != 0.0) acc += wi * a * (wt(ex) * wt(ey) * wt(ez));
                                    }
                                }
                            }
                        }
                    }
                }
            }
            gc[c * nc + q] = acc;
        }
    }
}

// ------------------------------------------------------------ transfers

// b_c = P^T r_f, unscaled full weighting (weights 1, 1/2 per dimension).
// A lane owns coarse column I = 1 + 64 bx + lane (the lane at I = 1 also
// owns I = 0) and marches K; on each fine row it reads the pair (2I-1, 2I)
// and takes 2I+1 from the next lane.  The in-plane weighted sum of the fine
// plane 2K+1 is reused as plane 2(K+1)-1.  Loops are uniform per wave.
// Grid: (ceil((mxc-1)/64), ceil(myc/TY) [3D] | ceil(ceil(myc/kch)/TY) [2D] | 1,
//        ceil(mzc/kch) [3D] | 1)
template <typename T, int DIM>
__device__ __forceinline__ void restrict_body(int bx, int by, int bz, int mxf, int myf, int mzf, int mxc,
                                              int myc, int mzc, int kch, const T* __restrict__ rf,
                                              T* __restrict__ bc) {
    const int lane = threadIdx.x;
    const int I = 1 + bx * TX + lane;
    const bool act = I < mxc;
    const bool own0 = I == 1;
    const int f = 2 * I;
    const bool he = f + 1 < mxf;
    const T h = T(0.5);
    // x-weighted sums on a fine row: a = coarse I, b = coarse 0 (lane I = 1)
    auto row = [&](long rb) -> T2<T> {
        T2<T> v{T(0), T(0)};
        if (act) v = ld2(rf + rb + f - 1);
        T nx = __shfl_down(v.a, 1, TX);
        if (act && lane == TX - 1 && he) nx = rf[rb + f + 1];
        T2<T> o;
        o.a = v.b + h * (v.a + (he ? nx : T(0)));
        o.b = own0 ? rf[rb] + h * rf[rb + 1] : T(0);
        return o;
    };
    auto comb = [&](T2<T> mid, T2<T> lo, T2<T> hi) -> T2<T> {
        return T2<T>{mid.a + h * (lo.a + hi.a), mid.b + h * (lo.b + hi.b)};
    };
    auto store = [&](long cb, T2<T> v) {
        if (act) bc[cb + I] = v.a;
        if (own0) bc[cb] = v.b;
    };
    const T2<T> zz{T(0), T(0)};
    if (DIM == 1) {
        store(0, row(0));
        return;
    }
    if (DIM == 2) {
        const int chunk = by * TY + threadIdx.y;
        const int J0 = chunk * kch, J1 = min(J0 + kch, myc);
        if (J0 >= J1) return;
        T2<T> slo = (2 * J0 - 1 >= 0) ? row((long)(2 * J0 - 1) * mxf) : zz;
        for (int J = J0; J < J1; ++J) {
            const int j = 2 * J;
            const T2<T> smid = row((long)j * mxf);
            const T2<T> shi = (j + 1 < myf) ? row((long)(j + 1) * mxf) : zz;
            store((long)J * mxc, comb(smid, slo, shi));
            slo = shi;
        }
        return;
    }
    // 3D: march the fine planes 2 K0 - 1 .. 2 K1 - 1 of the chunk; the three
    // rows of plane k+1 are loaded while plane k is summed
    const int J = by * TY + threadIdx.y;
    if (J >= myc) return;
    const int K0 = bz * kch, K1 = min(K0 + kch, mzc);
    if (K0 >= K1) return;
    const int j0 = 2 * J;
    const long sxy = (long)mxf * myf;
    const bool hlo = j0 - 1 >= 0, hhi = j0 + 1 < myf;
    // a fine row as loaded: the lane's pair (f-1, f) and one extra value
    // (last lane: f+1; the lane at I = 1: fine 0, with fine 1 = its pair's first)
    struct RR {
        T2<T> v;
        T x;
    };
    auto rload = [&](long rb) -> RR {
        RR r{zz, T(0)};
        if (act) {
            r.v = ld2(rf + rb + f - 1);
            if (own0) r.x = rf[rb];
            else if (lane == TX - 1 && he) r.x = rf[rb + f + 1];
        }
        return r;
    };
    auto rsum = [&](const RR& r) -> T2<T> {   // every lane of the wave takes part
        T nx = __shfl_down(r.v.a, 1, TX);
        if (act && lane == TX - 1 && he) nx = r.x;
        T2<T> o;
        o.a = r.v.b + h * (r.v.a + (he ? nx : T(0)));
        o.b = own0 ? r.x + h * r.v.a : T(0);
        return o;
    };
    RR c0{zz, T(0)}, c1{zz, T(0)}, c2{zz, T(0)}, n0{zz, T(0)}, n1{zz, T(0)}, n2{zz, T(0)};
    auto pload = [&](int k, RR& lo, RR& mid, RR& hi) {
        const long rb = (long)k * sxy + (long)j0 * mxf;
        mid = rload(rb);
        if (hlo) lo = rload(rb - mxf);
        if (hhi) hi = rload(rb + mxf);
    };
    const int kf0 = max(2 * K0 - 1, 0), kf1 = min(2 * K1 - 1, mzf - 1);
    pload(kf0, c0, c1, c2);
    T2<T> slo = zz, smid = zz;
    for (int k = kf0; k <= kf1; ++k) {
        if (k + 1 <= kf1) pload(k + 1, n0, n1, n2);
        const T2<T> s = comb(rsum(c1), hlo ? rsum(c0) : zz, hhi ? rsum(c2) : zz);
        if (k & 1) {                       // plane 2K+1: the last of K, the first of K+1
            if (k > 2 * K0) store(((long)(k >> 1) * myc + J) * mxc, comb(smid, slo, s));
            slo = s;
        } else {
            smid = s;
        }
        c0 = n0;
        c1 = n1;
        c2 = n2;
    }
    if (!(kf1 & 1)) store(((long)(kf1 >> 1) * myc + J) * mxc, comb(smid, slo, zz));   // the last fine plane
}

template <typename T, int DIM>
__global__ void __launch_bounds__(NT)
k_restrict(int mxf, int myf, int mzf, int mxc, int myc, int mzc, int kch, const T* __restrict__ rf,
           T* __restrict__ bc) {
    const Tile tl = xcd_tile();
    restrict_body<T, DIM>(tl.bx, tl.by, tl.bz, mxf, myf, mzf, mxc, myc, mzc, kch, rf, bc);
}

// b_c = P^T (b - A x) on a 3D constant-coefficient level without storing
// the fine residual: restrict_body applied to the residual that the pass
// stencil(SrcArr x, OutResid b) would store.  A block of TYC waves covers
// the coarse rows J0 .. J0+TYC-1 (J0 = TYC by) and the coarse columns
// I = 1 + 32 V bx + V/2 lane .. + V/2 - 1 (the lane at I = 1 also I = 0) of
// coarse planes K0 .. K1-1 (kch of them), marching the fine planes 2K0-1 ..
// 2K1-1.  Wave J forms the residual on the fine rows 2J-1, 2J, 2J+1 at the
// lane's V fine points 2I-1 .. 2I+V-2 (the wave's last lane also at the
// next one, the lane at I = 1 at fine 0) from the operand's rows 2J0-2 ..
// 2J0+2TYC, staged in LDS by LDS-DMA loads one plane ahead with the x-halo
// values is-1, is+W and is+W+1 (is = 1 + W bx, W = 64 V), and b read
// directly one plane ahead.  V = 2 (FP64) or 4 (FP32): 16-byte accesses.
// The residual at a point is OutResid's (the same operand, neighbours and
// operation order as stencil_vec) and the sums are restrict_body's: b_c
// equals the two separate passes up to rounding (the same formulas, whose
// FMA contractions the compiler chooses per kernel; FP64 cycles differ by
// about 2e-16 relative).
// Requires (mxf - 1) % W == 0.  Grid ((mxf-1)/W, ceil(myc/TYC),
// ceil(mzc/kch)), block (64, TYC).
template <typename T, int V, class Op, int TYC>
__device__ __forceinline__ void resid_restrict_lds(int bx, int by, int bz, int mx, int my, int mz, int mxc, int myc,
                                                   int mzc, int kch, const Op& op, const T* __restrict__ x,
                                                   const T* __restrict__ b, T* __restrict__ bc) {
    static_assert(V == 2 || V == 4, "two or four fine points per lane");
    constexpr int W = TX * V, VC = V / 2, NR = 2 * TYC + 3, S = 4;
    constexpr int DPE = sizeof(T) / 4, NLD = W * (int)sizeof(T) / 16, EPL = 16 / (int)sizeof(T);
    static_assert(DPE == 1 || DPE == 2, "4- or 8-byte elements");
    __shared__ __attribute__((aligned(16))) T ring[S][NR][W];
    __shared__ __attribute__((aligned(16))) T halo[S][NR][4];   // is-1, is+W, is+W+1
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int is = 1 + W * bx;
    const int J0 = TYC * by, rbase = 2 * J0 - 2;              // first staged fine row
    const int K0 = bz * kch, K1 = min(K0 + kch, mzc);
    if (K0 >= K1) return;                                      // uniform per block
    const int kf0 = max(2 * K0 - 1, 0), kf1 = min(2 * K1 - 1, mz - 1);
    const bool has_e = is + W < mx;                            // then is + W + 1 < mx too (full tiles)
    const long sm = (long)mx * my;
    const T h = T(0.5), z0 = T(0);
    auto stage_row = [&](int kk, int ro) {
        const int jr = rbase + ro;
        if (jr < 0 || jr >= my) return;
        const int slot = (kk + 4) % S;
        const T* g = x + (long)kk * sm + (long)jr * mx;
        if (lane < NLD)
            __builtin_amdgcn_global_load_lds((const void*)(g + is + EPL * lane), (lds_void_t*)&ring[slot][ro][0], 16,
                                             0, 0);
        if (lane < 3 * DPE) {
            const int e = lane / DPE;                          // 0: is-1, 1: is+W, 2: is+W+1
            if (e == 0 || has_e)
                __builtin_amdgcn_global_load_lds(
                    (const void*)((const char*)(g + (e == 0 ? is - 1 : is + W - 1 + e)) + 4 * (lane % DPE)),
                    (lds_void_t*)&halo[slot][ro][0], 4, 0, 0);
        }
    };
    auto stage = [&](int kk) {
        if (kk < kf0 - 1 || kk > kf1 + 1 || kk < 0 || kk >= mz) return;
        for (int ro = w; ro < NR; ro += TYC) stage_row(kk, ro);
    };
    const int J = J0 + w;
    const bool jok = J < myc;                                  // uniform per wave
    const int I = 1 + (W / 2) * bx + VC * lane;                // the lane's first coarse column
    const bool own0 = I == 1;
    const int i0 = 2 * I - 1;                                  // the lane's fine points i0 .. i0+V-1
    const bool he = i0 + V < mx;                               // fine i0+V exists
    TV<T, V> zv;
#pragma unroll
    for (int e = 0; e < V; ++e) zv.v[e] = z0;
    // b at the lane's points of fine row jr, plane kk; fine i0+V (last
    // lane); fine 0 (lane at I = 1)
    struct BV {
        TV<T, V> p;
        T e, z;
    };
    auto bload = [&](int kk, int jr) -> BV {
        BV v{zv, z0, z0};
        if (jr < 0 || jr >= my) return v;
        const long rq = (long)kk * sm + (long)jr * mx;
        v.p = ldv<T, V>(b + rq + i0);
        if (lane == TX - 1 && he) v.e = b[rq + i0 + V];
        if (own0) v.z = b[rq];
        return v;
    };
    // the residual at fine row jr (staged row ro) of plane kk (slots sd, sc, su)
    struct RV {
        TV<T, V> p;
        T e, z;
    };
    auto rl = [&](int sl, int ro) -> TV<T, V> { return ldv<T, V>(&ring[sl][ro][V * lane]); };
    auto resid_row = [&](int kk, int jr, int sd, int sc, int su, const BV& bb) -> RV {
        RV o{zv, z0, z0};
        if (jr < 0 || jr >= my) return o;
        const int ro = jr - rbase;
        const TV<T, V> uc = rl(sc, ro);
        const bool pb = jr == 0 || jr == my - 1 || kk == 0 || kk == mz - 1;   // boundary row or plane
        const bool ms = jr - 1 > 0, mn = jr + 1 < my - 1, md = kk - 1 > 0, mu = kk + 1 < mz - 1;
        TV<T, V> A;
        T d0;
        if (pb) {
#pragma unroll
            for (int e = 0; e < V; ++e) op.bnd(i0 + e, jr, kk, uc.v[e], A.v[e], d0);
        } else {
            const TV<T, V> sv = ms ? rl(sc, ro - 1) : zv;
            const TV<T, V> nv = mn ? rl(sc, ro + 1) : zv;
            const TV<T, V> dv = md ? rl(sd, ro) : zv;
            const TV<T, V> uv = mu ? rl(su, ro) : zv;
            const T wv = lane > 0 ? ring[sc][ro][V * lane - 1] : halo[sc][ro][0];
            const T ev = lane < TX - 1 ? ring[sc][ro][V * lane + V] : halo[sc][ro][1];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int ie = i0 + e;
                if (e > 0 && ie == mx - 1) {
                    op.bnd(ie, jr, kk, uc.v[e], A.v[e], d0);
                    continue;
                }
                const T we = e == 0 ? (ie - 1 > 0 ? wv : z0) : uc.v[e - 1];
                const T ee = ie + 1 < mx - 1 ? (e == V - 1 ? ev : uc.v[e + 1]) : z0;
                op.interior(ie, jr, kk, uc.v[e], we, ee, sv.v[e], nv.v[e], dv.v[e], uv.v[e], A.v[e], d0);
            }
            if (lane == TX - 1 && he) {                        // fine is+W: the next tile's first point
                const int ie = i0 + V;
                const T xe = halo[sc][ro][1];
                if (ie == mx - 1) {
                    op.bnd(ie, jr, kk, xe, o.e, d0);
                } else {
                    const T se = ms ? halo[sc][ro - 1][1] : z0, ne = mn ? halo[sc][ro + 1][1] : z0;
                    const T de = md ? halo[sd][ro][1] : z0, ue = mu ? halo[su][ro][1] : z0;
                    op.interior(ie, jr, kk, xe, ie - 1 > 0 ? uc.v[V - 1] : z0, ie + 1 < mx - 1 ? halo[sc][ro][2] : z0,
                                se, ne, de, ue, o.e, d0);
                }
                o.e = bb.e - o.e;
            }
        }
        if (pb && lane == TX - 1 && he) {
            T Ae;
            op.bnd(i0 + V, jr, kk, halo[sc][ro][1], Ae, d0);
            o.e = bb.e - Ae;
        }
#pragma unroll
        for (int e = 0; e < V; ++e) o.p.v[e] = bb.p.v[e] - A.v[e];
        if (own0) {                                            // fine 0: a boundary point
            T Az;
            op.bnd(0, jr, kk, halo[sc][ro][0], Az, d0);
            o.z = bb.z - Az;
        }
        return o;
    };
    // restrict_body's x-weighted sums of a fine row: a[m] = coarse I + m,
    // b = coarse 0
    struct CS {
        TV<T, VC> a;
        T b;
    };
    auto rsum = [&](const RV& v) -> CS {
        T nx = __shfl_down(v.p.v[0], 1, TX);
        if (lane == TX - 1) nx = v.e;
        CS o;
#pragma unroll
        for (int m = 0; m < VC; ++m) {
            const T right = 2 * m + 2 < V ? v.p.v[2 * m + 2] : (he ? nx : T(0));
            o.a.v[m] = v.p.v[2 * m + 1] + h * (v.p.v[2 * m] + right);
        }
        o.b = own0 ? v.z + h * v.p.v[0] : T(0);
        return o;
    };
    auto comb = [&](const CS& mid, const CS& lo, const CS& hi) -> CS {
        CS o;
#pragma unroll
        for (int m = 0; m < VC; ++m) o.a.v[m] = mid.a.v[m] + h * (lo.a.v[m] + hi.a.v[m]);
        o.b = mid.b + h * (lo.b + hi.b);
        return o;
    };
    auto store = [&](int K, const CS& v) {
        const long cb = ((long)K * myc + J) * mxc;
#pragma unroll
        for (int m = 0; m < VC; ++m)
            if (I + m < mxc) bc[cb + I + m] = v.a.v[m];
        if (own0) bc[cb] = v.b;
    };
    const bool hlo = 2 * J - 1 >= 0, hhi = 2 * J + 1 < my;
    BV cb0, cb1, cb2, nb0, nb1, nb2;
    if (jok) {
        cb0 = bload(kf0, 2 * J - 1);
        cb1 = bload(kf0, 2 * J);
        cb2 = bload(kf0, 2 * J + 1);
    }
    for (int kk = kf0 - 1; kk <= kf0 + 1; ++kk) stage(kk);
    CS zz;
#pragma unroll
    for (int m = 0; m < VC; ++m) zz.a.v[m] = z0;
    zz.b = z0;
    CS slo = zz, smid = zz;
    for (int k = kf0; k <= kf1; ++k) {
        wait_vmcnt(0);
        lds_barrier();
        stage(k + 2);
        if (!jok) continue;                                    // uniform per wave (no barrier skipped)
        if (k + 1 <= kf1) {
            nb0 = bload(k + 1, 2 * J - 1);
            nb1 = bload(k + 1, 2 * J);
            nb2 = bload(k + 1, 2 * J + 1);
        }
        const int sd = (k + 3) % S, sc = (k + 4) % S, su = (k + 5) % S;
        const RV r1 = resid_row(k, 2 * J, sd, sc, su, cb1);
        const RV r0 = resid_row(k, 2 * J - 1, sd, sc, su, cb0);
        const RV r2 = resid_row(k, 2 * J + 1, sd, sc, su, cb2);
        const CS sv = comb(rsum(r1), hlo ? rsum(r0) : zz, hhi ? rsum(r2) : zz);
        if (k & 1) {                                           // plane 2K+1: the last of K, the first of K+1
            if (k > 2 * K0) store(k >> 1, comb(smid, slo, sv));
            slo = sv;
        } else {
            smid = sv;
        }
        cb0 = nb0;
        cb1 = nb1;
        cb2 = nb2;
    }
    if (jok && !(kf1 & 1)) store(kf1 >> 1, comb(smid, slo, zz));   // the last fine plane
}

template <typename T, int V, class Op, int TYC>
__global__ void __launch_bounds__(TX * TYC)
k_resid_restrict_lds(int mx, int my, int mz, int mxc, int myc, int mzc, int kch, Op op, const T* __restrict__ x,
                     const T* __restrict__ b, T* __restrict__ bc) {
    const Tile tl = xcd_tile();
    resid_restrict_lds<T, V, Op, TYC>(tl.bx, tl.by, tl.bz, mx, my, mz, mxc, myc, mzc, kch, op, x, b, bc);
}

// out = base + P x_c, base = c b (ZG: zero-guess smoothing result) or x.
// A lane owns coarse column I = 1 + 64 bx + lane and writes the fine pair
// (2I-1, 2I) (the lane at I = 1 also writes fine 0); coarse I-1 comes from
// the previous lane.  Fine rows j (3D) march k; the coarse in-plane
// interpolant of plane K is computed once and reused for fine planes 2K-1,
// 2K, 2K+1.  Loops are uniform per wave.
// Grid: (ceil((mxc-1)/64), ceil(myf/TY) [3D] | ceil(ceil(myf/kch)/TY) [2D] | 1,
//        ceil(mzf/kch) [3D] | 1)
template <typename T, int DIM, bool ZG>
__device__ __forceinline__ void prolong_body(int bx, int by, int bz, int mxf, int myf, int mzf, int mxc,
                                             int myc, int kch, const T* __restrict__ xb, T c,
                                             const T* __restrict__ xc, T* __restrict__ out) {
    const int lane = threadIdx.x;
    const int I = 1 + bx * TX + lane;
    const bool act = I < mxc;
    const bool own0 = I == 1;
    const int f = 2 * I;
    const T h = T(0.5);
    // coarse row values: a = coarse I-1, b = coarse I
    auto crow = [&](long cb) -> T2<T> {
        const T v = act ? xc[cb + I] : T(0);
        T w = __shfl_up(v, 1, TX);
        if (act && lane == 0) w = xc[cb + I - 1];
        return T2<T>{w, v};
    };
    auto mix = [&](T2<T> lo, T2<T> hi) -> T2<T> { return T2<T>{h * (lo.a + hi.a), h * (lo.b + hi.b)}; };
    auto emit = [&](long fb, T2<T> v) {          // fine row base fb, coarse-row values v
        if (act) {
            const T2<T> bb = ld2(xb + fb + f - 1);
            const T b0 = ZG ? c * bb.a : bb.a, b1 = ZG ? c * bb.b : bb.b;
            st2(out + fb + f - 1, b0 + h * (v.a + v.b), b1 + v.b);
        }
        if (own0) out[fb] = (ZG ? c * xb[fb] : xb[fb]) + v.a;
    };
    if (DIM == 1) {
        emit(0, crow(0));
        return;
    }
    if (DIM == 2) {
        const int chunk = by * TY + threadIdx.y;
        const int j0 = chunk * kch, j1 = min(j0 + kch, myf);
        for (int j = j0; j < j1; ++j) {
            const int ja = j >> 1, jb = (j + 1) >> 1;
            T2<T> v = crow((long)ja * mxc);
            if (j & 1) v = mix(v, crow((long)jb * mxc));
            emit((long)j * mxf, v);
        }
        return;
    }
    const int j = by * TY + threadIdx.y;
    if (j >= myf) return;
    const int ja = j >> 1, jb = (j + 1) >> 1;
    const bool jodd = j & 1;
    const long sxyc = (long)mxc * myc;
    auto cplane = [&](int K) -> T2<T> {
        const long cb = (long)K * sxyc;
        const T2<T> v = crow(cb + (long)ja * mxc);
        return jodd ? mix(v, crow(cb + (long)jb * mxc)) : v;
    };
    const int k0 = bz * kch, k1 = min(k0 + kch, mzf);
    if (k0 >= k1) return;
    const long sxy = (long)mxf * myf;
    const T2<T> zz{T(0), T(0)};
    int K = k0 >> 1;
    T2<T> plo = cplane(K);
    T2<T> phi = (2 * K + 1 < mzf) ? cplane(K + 1) : zz;
    for (int k = k0; k < k1; ++k) {
        const int Kk = k >> 1;
        if (Kk != K) {     // advance one coarse plane
            K = Kk;
            plo = phi;
            phi = (2 * K + 1 < mzf) ? cplane(K + 1) : zz;
        }
        emit((long)k * sxy + (long)j * mxf, (k & 1) ? mix(plo, phi) : plo);
    }
}

template <typename T, int DIM, bool ZG>
__global__ void __launch_bounds__(NT)
k_prolong(int mxf, int myf, int mzf, int mxc, int myc, int kch, const T* __restrict__ xb, T c,
          const T* __restrict__ xc, T* __restrict__ out) {
    const Tile tl = xcd_tile();
    prolong_body<T, DIM, ZG>(tl.bx, tl.by, tl.bz, mxf, myf, mzf, mxc, myc, kch, xb, c, xc, out);
}

// ------------------------------------------- fused prolongation + smoothing

// Values of one coarse row seen by a lane: w = coarse I-1, v = coarse I,
// e = coarse I+1 (e is needed by the wave's last lane only).
template <typename T>
struct T3 {
    T w, v, e;
};

template <typename T>
__device__ __forceinline__ T3<T> mix3(const T3<T>& a, const T3<T>& b) {
    const T h = T(0.5);
    return T3<T>{h * (a.w + b.w), h * (a.v + b.v), h * (a.e + b.e)};
}

// One smoothing step applied to the prolonged iterate without storing it:
//   t = base + P x_c   (base = c xb with ZG, else xb)
//   Out(t, A t)        (e.g. y = t + s D^-1 (b - A t))
// on a 2D/3D level of size (mx, my, mz) with coarse level (mxc, myc, .).
// A lane owns coarse column I = 1 + 64 bx + lane and the fine pair
// (2I-1, 2I) -- the lane mapping of stencil_vec with V = 2 -- and the lane at I = 1
// also fine 0.  The coarse values of I-1 and I+1 come from the neighbouring
// lanes; t on a fine row is interpolated from the coarse rows of the two
// coarse planes k>>1 and (k>>1)+1.
//
// 3D form of the fused step.  A block of R+2 waves covers the R fine rows
// j0 .. j0+R-1 (j0 = R by) and planes k0 .. k1-1 (k0 = kch bz); wave w forms
// t on row j0-1+w: the first and last wave form only the rows beside the
// block (read by the stencil of its edge rows), the others also apply the
// stencil to their row.  t of plane k goes through LDS (double-buffered by
// plane parity, one barrier per plane; every wave runs the same loop); t
// of planes k-1, k, k+1 of the own row stays in registers.  Loads run ahead
// of use: the base row two planes ahead, the coarse row of the next coarse
// plane one coarse plane ahead.  Grid (ceil((mx-1)/128), ceil(my/R),
// ceil(mz/kch)), block (64, R+2).
template <typename T, class Op, bool ZG, class Out, int R>
__device__ __forceinline__ void prolong_smooth_3d(int bx, int by, int bz, int mx, int my, int mz, int mxc,
                                                  int myc, int kch, const Op& op, const T* __restrict__ xb,
                                                  T c, const T* __restrict__ xc, const Out& out, double& a0,
                                                  double& a1) {
    __shared__ T2<T> tbuf[2][R + 2][TX];
    const int lane = threadIdx.x, w = threadIdx.y;
    const int I = 1 + bx * TX + lane;
    const bool act = I < mxc;
    const int i0 = 2 * I - 1, i1 = 2 * I;
    const T h = T(0.5), z0 = T(0);
    const T3<T> zero3{z0, z0, z0};
    const T2<T> zero2{z0, z0};
    const int j = by * R + w - 1;
    const bool halo = w == 0 || w == R + 1;
    // rows this wave forms t on: an output row inside the grid; a halo row
    // only where a stencil reads it (interior rows)
    const bool rowok = halo ? (j > 0 && j < my - 1) : (j < my);   // uniform per wave
    const int k0 = bz * kch, k1 = min(k0 + kch, mz);
    if (k0 >= k1) return;                // uniform per block
    const int mzc = (mz - 1) / 2 + 1;
    const long sy = mx, sz = (long)mx * my, scz = (long)mxc * myc;
    const bool la = act && rowok;
    const int J = rowok ? j >> 1 : 0;
    const bool odd = (j & 1) != 0;
    // coarse values of the own row on one coarse plane, as loaded: v = coarse
    // column I, x = the edge lane's extra column (lane 0: I-1, last lane:
    // I+1); second row (J+1) for odd j
    struct CR {
        T v0, x0, v1, x1;
    };
    const CR cr0{z0, z0, z0, z0};
    auto craw = [&](int K) -> CR {
        CR r = cr0;
        if (la) {
            const long pb = (long)K * scz + (long)J * mxc + I;
            r.v0 = xc[pb];
            if (lane == 0) r.x0 = xc[pb - 1];
            else if (lane == TX - 1 && I + 1 < mxc) r.x0 = xc[pb + 1];
            if (odd) {
                r.v1 = xc[pb + mxc];
                if (lane == 0) r.x1 = xc[pb + mxc - 1];
                else if (lane == TX - 1 && I + 1 < mxc) r.x1 = xc[pb + mxc + 1];
            }
        }
        return r;
    };
    // the row's coarse values (I-1, I, I+1) from a loaded CR (every lane of
    // the wave takes part in the exchange)
    auto fin1 = [&](T v, T x) -> T3<T> {
        T wv = __shfl_up(v, 1, TX);
        T ev = __shfl_down(v, 1, TX);
        if (act && lane == 0) wv = x;
        if (act && lane == TX - 1 && I + 1 < mxc) ev = x;
        return T3<T>{wv, v, ev};
    };
    auto cfin = [&](const CR& r) -> T3<T> {
        const T3<T> a = fin1(r.v0, r.x0);
        if (!odd) return a;
        return mix3(a, fin1(r.v1, r.x1));
    };
    struct BP {
        T2<T> p;
        T x;
    };
    // base values on the own row: the lane's pair and the edge lane's extra value
    auto bload = [&](long qq) -> BP {
        BP r{zero2, z0};
        if (la) {
            r.p = ld2(xb + qq);
            if (lane == 0) r.x = xb[qq - 1];
            else if (lane == TX - 1 && i1 + 1 < mx) r.x = xb[qq + 2];
        }
        return r;
    };
    auto tpb = [&](const T3<T>& P, const T2<T>& bb) -> T2<T> {
        const T b0 = ZG ? c * bb.a : bb.a, b1 = ZG ? c * bb.b : bb.b;
        return T2<T>{b0 + h * (P.w + P.v), b1 + P.v};
    };
    auto scb = [&](T v) -> T { return ZG ? c * v : v; };
    int Ka = k0 >> 1;
    long q = (long)k0 * sz + (long)(rowok ? j : 0) * sy + i0;
    T3<T> Ya = zero3, Yb = zero3, Pc = zero3;
    CR cn = cr0;
    T2<T> tc = zero2, td = zero2;
    BP bp0{zero2, z0}, bp1{zero2, z0};
    if (rowok) {
        Ya = cfin(craw(Ka));
        if (Ka + 1 < mzc) Yb = cfin(craw(Ka + 1));
        if (Ka + 2 < mzc) cn = craw(Ka + 2);
        Pc = (k0 & 1) ? mix3(Ya, Yb) : Ya;
        bp0 = bload(q);
        if (k0 + 1 < mz) bp1 = bload(q + sz);
        tc = tpb(Pc, bp0.p);
        if (!halo && k0 > 0) {
            const T3<T> Pd = (k0 & 1) ? Ya : mix3(cfin(craw(Ka - 1)), Ya);
            const BP bd = bload(q - sz);
            td = tpb(Pd, bd.p);
        }
    }
    const bool jb = (j == 0) || (j == my - 1);
    for (int k = k0; k < k1; ++k, q += sz) {
        T3<T> Pu = zero3;
        T2<T> tu = zero2;
        BP bp2{zero2, z0};
        if (rowok) {
            if (k + 2 < mz) bp2 = bload(q + 2 * sz);
            if ((k >> 1) != Ka) {
                Ka = k >> 1;
                Ya = Yb;
                Yb = (Ka + 1 < mzc) ? cfin(cn) : zero3;
                cn = (Ka + 2 < mzc) ? craw(Ka + 2) : cr0;
            }
            if (k + 1 < mz) {
                Pu = (k & 1) ? Yb : mix3(Ya, Yb);
                tu = tpb(Pu, bp1.p);
            }
            if (act) tbuf[k & 1][w][lane] = tc;
        }
        __syncthreads();
        T wv = z0, ev = z0;
        if (!halo && rowok) {            // uniform per wave: the whole wave exchanges
            wv = __shfl_up(tc.b, 1, TX);
            ev = __shfl_down(tc.a, 1, TX);
        }
        if (!halo && rowok && act) {
            const bool need_sn = !jb && k != 0 && k != mz - 1;
            const T2<T> ts = (need_sn && j - 1 > 0) ? tbuf[k & 1][w - 1][lane] : zero2;
            const T2<T> tn = (need_sn && j + 1 < my - 1) ? tbuf[k & 1][w + 1][lane] : zero2;
            const T2<T> tdd = (k - 1 > 0) ? td : zero2, tuu = (k + 1 < mz - 1) ? tu : zero2;
            if (lane == 0 && i0 - 1 > 0) wv = scb(bp0.x) + Pc.w;
            if (lane == TX - 1 && i1 + 1 < mx - 1) ev = scb(bp0.x) + h * (Pc.v + Pc.e);
            T Aa, Ab, ia, ib;
            if (jb || k == 0 || k == mz - 1) {
                op.bnd(i0, j, k, tc.a, Aa, ia);
                op.bnd(i1, j, k, tc.b, Ab, ib);
            } else {
                op.interior(i0, j, k, tc.a, i0 - 1 > 0 ? wv : z0, i1 < mx - 1 ? tc.b : z0, ts.a, tn.a, tdd.a, tuu.a,
                            Aa, ia);
                if (i1 == mx - 1) op.bnd(i1, j, k, tc.b, Ab, ib);
                else op.interior(i1, j, k, tc.b, tc.a, i1 + 1 < mx - 1 ? ev : z0, ts.b, tn.b, tdd.b, tuu.b, Ab, ib);
            }
            const TV<T, 2> uu{{tc.a, tc.b}}, AA{{Aa, Ab}}, II{{ia, ib}};
            // with ZG the base is the output's b, its one per-point array
            if constexpr (ZG) {
                static_assert(Out::NPRE == 1, "the base is the output's b");
                const TV<T, 2> bb{{bp0.p.a, bp0.p.b}};
                out.template pointvp<2>(q, uu, AA, II, &bb, a0, a1);
            } else {
                out.template pointv<2>(q, uu, AA, II, a0, a1);
            }
            if (I == 1) {      // fine 0: a boundary point
                const T t0 = scb(bp0.x) + Pc.w;
                op.bnd(0, j, k, t0, Aa, ia);
                if constexpr (ZG) out.pointp(q - 1, t0, Aa, ia, &bp0.x, a0, a1);
                else out.point(q - 1, t0, Aa, ia, a0, a1);
            }
        }
        td = tc;
        tc = tu;
        Pc = Pu;
        bp0 = bp1;
        bp1 = bp2;
    }
}

template <typename T, int DIM, class Op, bool ZG, class Out>
__device__ __forceinline__ void prolong_smooth_body(int bx, int by, int bz, int mx, int my, int mz, int mxc,
                                                    int myc, int kch, const Op& op, const T* __restrict__ xb,
                                                    T c, const T* __restrict__ xc, const Out& out, double& a0,
                                                    double& a1) {
    // 2D: a thread marches fine rows (grid: that of stencil_vec with V = 2)
    if constexpr (DIM == 3) {
        prolong_smooth_3d<T, Op, ZG, Out, TYF>(bx, by, bz, mx, my, mz, mxc, myc, kch, op, xb, c, xc, out, a0, a1);
        return;
    }
    const int lane = threadIdx.x;
    const int I = 1 + bx * TX + lane;
    const bool act = I < mxc;
    const int i0 = 2 * I - 1, i1 = 2 * I;
    const T h = T(0.5), z0 = T(0);
    const T3<T> zero3{z0, z0, z0};
    const T2<T> zero2{z0, z0};
    auto crow = [&](long cb) -> T3<T> {
        const T v = act ? xc[cb + I] : z0;
        T w = __shfl_up(v, 1, TX);
        T e = __shfl_down(v, 1, TX);
        if (act && lane == 0) w = xc[cb + I - 1];
        if (act && lane == TX - 1 && I + 1 < mxc) e = xc[cb + I + 1];
        return T3<T>{w, v, e};
    };
    auto basev = [&](long q) -> T { return ZG ? c * xb[q] : xb[q]; };
    auto tpair = [&](const T3<T>& P, long q) -> T2<T> {   // t at (i0, i1), address q of i0
        const T2<T> bb = ld2(xb + q);
        const T b0 = ZG ? c * bb.a : bb.a, b1 = ZG ? c * bb.b : bb.b;
        return T2<T>{b0 + h * (P.w + P.v), b1 + P.v};
    };
    // the stencil and output at the pair of row j, march index m
    auto emit = [&](long q, int j, int m, int M, bool rowb, const T3<T>& Pc, const T2<T>& tc, const T2<T>& ts,
                    const T2<T>& tn, const T2<T>& td, const T2<T>& tu, T wv, T ev) {
        const int jj = DIM == 3 ? j : m, kk = DIM == 3 ? m : 0;
        if (lane == 0 && i0 - 1 > 0) wv = basev(q - 1) + Pc.w;
        if (lane == TX - 1 && i1 + 1 < mx - 1) ev = basev(q + 2) + h * (Pc.v + Pc.e);
        T Aa, Ab, ia, ib;
        if (rowb || m == 0 || m == M - 1) {
            op.bnd(i0, jj, kk, tc.a, Aa, ia);
            op.bnd(i1, jj, kk, tc.b, Ab, ib);
        } else {
            op.interior(i0, jj, kk, tc.a, i0 - 1 > 0 ? wv : z0, i1 < mx - 1 ? tc.b : z0, ts.a, tn.a, td.a, tu.a,
                        Aa, ia);
            if (i1 == mx - 1) op.bnd(i1, jj, kk, tc.b, Ab, ib);
            else op.interior(i1, jj, kk, tc.b, tc.a, i1 + 1 < mx - 1 ? ev : z0, ts.b, tn.b, td.b, tu.b, Ab, ib);
        }
        out.template pointv<2>(q, TV<T, 2>{{tc.a, tc.b}}, TV<T, 2>{{Aa, Ab}}, TV<T, 2>{{ia, ib}}, a0, a1);
        if (I == 1) {      // fine 0: a boundary point
            const T t0 = basev(q - 1) + Pc.w;
            op.bnd(0, jj, kk, t0, Aa, ia);
            out.point(q - 1, t0, Aa, ia, a0, a1);
        }
    };
    if (DIM == 2) {
        const int m0 = (by * TY + threadIdx.y) * kch, m1 = min(m0 + kch, my);
        if (m0 >= m1) return;
        const long sm = mx;
        int Ka = m0 >> 1;
        T3<T> Ya = crow((long)Ka * mxc);
        T3<T> Yb = (Ka + 1 < myc) ? crow((long)(Ka + 1) * mxc) : zero3;
        long q = (long)m0 * sm + i0;
        T3<T> Pc = (m0 & 1) ? mix3(Ya, Yb) : Ya;
        T2<T> tc = act ? tpair(Pc, q) : zero2, td = zero2;
        if (m0 > 0) {
            const T3<T> Pd = (m0 & 1) ? Ya : mix3(crow((long)(Ka - 1) * mxc), Ya);
            if (act) td = tpair(Pd, q - sm);
        }
        for (int m = m0; m < m1; ++m, q += sm) {
            if ((m >> 1) != Ka) {
                Ka = m >> 1;
                Ya = Yb;
                Yb = (Ka + 1 < myc) ? crow((long)(Ka + 1) * mxc) : zero3;
            }
            T3<T> Pu = zero3;
            T2<T> tu = zero2;
            if (m + 1 < my) {
                Pu = (m & 1) ? Yb : mix3(Ya, Yb);
                if (act) tu = tpair(Pu, q + sm);
            }
            const T wv = __shfl_up(tc.b, 1, TX), ev = __shfl_down(tc.a, 1, TX);
            if (act)
                emit(q, 0, m, my, false, Pc, tc, (m - 1 > 0) ? td : zero2, (m + 1 < my - 1) ? tu : zero2, zero2,
                     zero2, wv, ev);
            td = tc;
            tc = tu;
            Pc = Pu;
        }
        return;
    }
}

template <typename T, int DIM, class Op, bool ZG, class Out>
__global__ void __launch_bounds__(NT)
k_prolong_smooth(int mx, int my, int mz, int mxc, int myc, int kch, Op op, const T* __restrict__ xb, T c,
                 const T* __restrict__ xc, Out out, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    const Tile tl = xcd_tile();
    prolong_smooth_body<T, DIM, Op, ZG, Out>(tl.bx, tl.by, tl.bz, mx, my, mz, mxc, myc, kch, op, xb, c, xc, out,
                                             a0, a1);
    block_partials<Out::NDOT>(a0, a1, part, tl.id);
}

// ----------------------------- LDS-staged fused prolongation + smoothing
//
// prolong_smooth_3d (one smoothing step applied to t = base + P x_c without
// storing t) with its inputs staged in LDS, V fine points per lane (FP64: 2,
// FP32: 4, 16-byte accesses either way).
// A block of TYD waves covers the fine rows j0 .. j0+TYD-1 (j0 = TYD by) of
// W = 64 V points (is = 1 + W bx, coarse columns W/2 bx .. W/2 bx + W/2 + 1) and
// marches the planes k0 .. k1-1.  LDS-DMA loads stage each base plane's
// rows j0-1 .. j0+TYD (and their x-halo values) one plane ahead, and each
// coarse plane's rows (j0-1)/2 .. (j0+TYD)/2 when the first fine plane that
// interpolates from it is staged.  At step k the waves form t on plane k+1
// (rows j0-1 .. j0+TYD, points is-1 .. is+W) into a ring of three t planes,
// with prolong_smooth_3d's arithmetic (rows mixed, then planes, then x), a
// barrier makes it visible, and each wave applies the stencil to its row of
// plane k from t on planes k-1 .. k+1, as stencil_lds does.  Every wave
// waits for all its loads at each step (s_waitcnt vmcnt(0)), so loads of
// the next plane overlap one step of arithmetic.  Base slots are refilled
// two steps after their last read, t and coarse slots likewise, each behind
// a barrier.  Requires (mx - 1) % W == 0; the stencil's per-point order is
// prolong_smooth_3d's, the block partition (and so the dot-product
// summation order) is stencil_lds's.
template <typename T, int V, class Op, bool ZG, class Out, int TYD>
__device__ __forceinline__ void prolong_smooth_lds(int bx, int by, int bz, int mx, int my, int mz, int mxc, int myc,
                                                   int kch, const Op& op, const T* __restrict__ xb, T c,
                                                   const T* __restrict__ xc, const Out& out, double& a0,
                                                   double& a1) {
    static_assert(V == 2 || V == 4, "two or four fine points per lane");
    constexpr int W = TX * V, CW = W / 2 + 2, NCR = TYD / 2 + 2, VC = V / 2;   // VC coarse intervals per lane
    // LDS-DMA geometry: dwords per element, elements per 16-byte lane load,
    // lanes that load a W-point row and a W/2-point coarse row
    constexpr int DPE = (int)sizeof(T) / 4, EPL = 16 / (int)sizeof(T);
    constexpr int NLD = W * (int)sizeof(T) / 16, NLDC = (W / 2) * (int)sizeof(T) / 16;
    static_assert(DPE == 1 || DPE == 2, "4- or 8-byte elements");
    constexpr int SB = 3, ST = 3, SC = 4;
    constexpr int NP = ZG ? 0 : Out::NPRE, NPA = NP > 0 ? NP : 1;
    static_assert(!ZG || Out::NPRE == 1, "with ZG the base is the output's b");
    __shared__ __attribute__((aligned(16))) T bring[SB][TYD + 2][W];     // base rows
    __shared__ __attribute__((aligned(16))) T bhalo[SB][TYD + 2][2];     //   and x-halo values
    __shared__ __attribute__((aligned(16))) T tring[ST][TYD + 2][W];     // t rows
    __shared__ __attribute__((aligned(16))) T thalo[ST][TYD + 2][2];
    __shared__ __attribute__((aligned(16))) T cring[SC][NCR][CW + 2];    // coarse rows
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int is = 1 + W * bx, ic = (W / 2) * bx;    // first fine point, first coarse column
    const int j0 = TYD * by;
    const int k0 = bz * kch, k1 = min(k0 + kch, mz);
    if (k0 >= k1) return;                            // uniform per block
    const int mzc = (mz - 1) / 2 + 1;
    const int klast = min(k1, mz - 1);               // last plane t is formed on
    const int Jlo = max(0, (j0 - 1) >> 1);           // first staged coarse row
    const long sm = (long)mx * my, smc = (long)mxc * myc;
    const bool has_e = is + W < mx;                  // fine is+W exists (then so does coarse ic+65)
    const T h = T(0.5), z0 = T(0);
    // base plane kk: rows w and w + TYD of this wave (rows j0 - 1 + r)
    auto stage_base_row = [&](int kk, int r) {
        const int jr = j0 - 1 + r;
        if (jr < 0 || jr >= my) return;
        const int slot = (kk + 3) % SB;
        const T* g = xb + (long)kk * sm + (long)jr * mx;
        if (lane < NLD)
            __builtin_amdgcn_global_load_lds((const void*)(g + is + EPL * lane), (lds_void_t*)&bring[slot][r][0], 16, 0,
                                             0);
        if (lane < 2 * DPE) {   // dword lane % DPE of the west (is-1) or east (is+W) value
            const bool east = lane >= DPE;
            if (!east || has_e)
                __builtin_amdgcn_global_load_lds(
                    (const void*)((const char*)(g + (east ? is + W : is - 1)) + 4 * (lane % DPE)),
                    (lds_void_t*)&bhalo[slot][r][0], 4, 0, 0);
        }
    };
    // coarse plane K: row Jlo + w of this wave (w < NCR), columns ic .. ic+65
    auto stage_coarse = [&](int K) {
        if (K < 0 || K >= mzc || w >= NCR) return;
        const int J = Jlo + w;
        if (J >= myc) return;
        const T* g = xc + (long)K * smc + (long)J * mxc + ic;
        if (lane < NLDC)
            __builtin_amdgcn_global_load_lds((const void*)(g + EPL * lane), (lds_void_t*)&cring[K % SC][w][0], 16, 0,
                                             0);
        if (lane < 2 * DPE) {   // columns ic+W/2 and ic+W/2+1 (in range when fine is+W is), a dword per lane
            if (lane < DPE || has_e)
                __builtin_amdgcn_global_load_lds(
                    (const void*)((const char*)(g + W / 2 + lane / DPE) + 4 * (lane % DPE)),
                    (lds_void_t*)&cring[K % SC][w][W / 2], 4, 0, 0);
        }
    };
    int kst = -1;   // highest coarse plane staged (uniform)
    auto stage = [&](int kk) {
        if (kk < 0 || kk > klast) return;
        stage_base_row(kk, w);
        if (w + TYD < TYD + 2) stage_base_row(kk, w + TYD);
        const int kn = (kk >> 1) + (kk & 1);   // the coarse planes kk interpolates from: kk/2 .. kn
        if (kst < (kk >> 1) - 1) kst = (kk >> 1) - 1;
        while (kst < kn) stage_coarse(++kst);
    };
    // t on plane kk (rows j0-1 .. j0+TYD) into its ring slot: rows w, w + TYD
    auto form_row = [&](int kk, int r) {
        const int j = j0 - 1 + r;
        if (j < 0 || j >= my) return;
        const int bs = (kk + 3) % SB, ts = (kk + 3) % ST;
        const int Ka = kk >> 1, Kb = Ka + (kk & 1);
        const int ra = (j >> 1) - Jlo, rb = ra + (j & 1);
        // P x_c at local coarse column cl of this row and plane (rows, then planes)
        auto pc = [&](int cl) -> T {
            const T* ca = &cring[Ka % SC][0][0];
            const T* cb = &cring[Kb % SC][0][0];
            T ya = ca[ra * (CW + 2) + cl], yb;
            if (j & 1) ya = h * (ya + ca[rb * (CW + 2) + cl]);
            if (kk & 1) {
                yb = cb[ra * (CW + 2) + cl];
                if (j & 1) yb = h * (yb + cb[rb * (CW + 2) + cl]);
                ya = h * (ya + yb);
            }
            return ya;
        };
        // coarse local columns VC lane .. VC lane + VC: fine point e of the
        // lane is odd (between two of them) for even e, on one for odd e
        T pcl[VC + 1];
#pragma unroll
        for (int m = 0; m <= VC; ++m) pcl[m] = pc(VC * lane + m);
        const TV<T, V> bb = ldv<T, V>(&bring[bs][r][V * lane]);
        TV<T, V> tt;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const T be = ZG ? c * bb.v[e] : bb.v[e];
            tt.v[e] = (e & 1) ? be + pcl[(e + 1) / 2] : be + h * (pcl[e / 2] + pcl[e / 2 + 1]);
        }
        stv<T, V>(&tring[ts][r][V * lane], tt);
        if (lane == 0) {                                 // fine is-1 (even): coarse ic
            const T x = bhalo[bs][r][0];
            thalo[ts][r][0] = (ZG ? c * x : x) + pcl[0];
        }
        if (lane == TX - 1 && has_e) {                   // fine is+W (odd): coarse ic+W/2, ic+W/2+1
            const T x = bhalo[bs][r][1];
            thalo[ts][r][1] = (ZG ? c * x : x) + h * (pcl[VC] + pc(VC * lane + VC + 1));
        }
    };
    const int j = j0 + w;
    const bool rowok = j < my;                       // uniform per wave
    const bool jb = j == 0 || j == my - 1;
    const bool ms = j - 1 > 0, mn = j + 1 < my - 1;
    const int r = w + 1;
    const int i0 = is + V * lane;
    TV<T, V> zv;
#pragma unroll
    for (int e = 0; e < V; ++e) zv.v[e] = z0;
    TV<T, V> co[NPA], ro[NPA];
    T cox[NPA], rox[NPA];
#pragma unroll
    for (int q = 0; q < NPA; ++q) {
        co[q] = ro[q] = zv;
        cox[q] = rox[q] = z0;
    }
    auto pre_own = [&](long qq, TV<T, V>* o, T* ox) {
        if constexpr (NP > 0) {
            out.template pre<V>(qq, o);
            if (i0 == 1) out.pre1(qq - 1, ox);
        }
    };
    auto tv = [&](int ts, int rr) -> TV<T, V> { return ldv<T, V>(&tring[ts][rr][V * lane]); };
    long q = (long)k0 * sm + (long)j * mx + i0;
    if (rowok) pre_own(q, co, cox);
    TV<T, V> ud = zv, uc = zv;
    stage(k0 - 1);
    for (int k = k0 - 2; k < k1; ++k) {
        wait_vmcnt(0);
        lds_barrier();
        stage(k + 2);
        if (k + 1 >= 0 && k + 1 >= k0 - 1 && k + 1 <= klast) {
            form_row(k + 1, w);
            if (w + TYD < TYD + 2) form_row(k + 1, w + TYD);
        }
        lds_barrier();
        if (k < k0 || !rowok) continue;             // uniform per wave (no barrier skipped)
        if (k + 1 < k1) pre_own(q + sm, ro, rox);
        const int td = (k + 2) % ST, tc = (k + 3) % ST, tu = (k + 4) % ST;
        if (k == k0) {
            uc = tv(tc, r);
            if (k > 0) ud = tv(td, r);
        }
        const TV<T, V> uu = k + 1 < mz ? tv(tu, r) : zv;
        TV<T, V> Au = zv, idg = zv;
        if (jb || k == 0 || k == mz - 1) {
#pragma unroll
            for (int e = 0; e < V; ++e) op.bnd(i0 + e, j, k, uc.v[e], Au.v[e], idg.v[e]);
        } else {
            // neighbours that are boundary points are left out (elementwise
            // selects: a select of whole vectors can make the compiler
            // address them through scratch memory)
            const TV<T, V> sv = ms ? tv(tc, r - 1) : zv, nv = mn ? tv(tc, r + 1) : zv;
            const bool hd = k - 1 > 0, hu = k + 1 < mz - 1;
            TV<T, V> dn, up;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                dn.v[e] = hd ? ud.v[e] : z0;
                up.v[e] = hu ? uu.v[e] : z0;
            }
            const T wv = lane > 0 ? tring[tc][r][V * lane - 1] : thalo[tc][r][0];
            const T ev = lane < TX - 1 ? tring[tc][r][V * lane + V] : thalo[tc][r][1];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int ie = i0 + e;
                if (e > 0 && ie == mx - 1) {             // the last point of the last tile: boundary
                    op.bnd(ie, j, k, uc.v[e], Au.v[e], idg.v[e]);
                    continue;
                }
                const T we = e == 0 ? (ie - 1 > 0 ? wv : z0) : uc.v[e - 1];
                const T ee = ie + 1 < mx - 1 ? (e == V - 1 ? ev : uc.v[e + 1]) : z0;
                op.interior(ie, j, k, uc.v[e], we, ee, sv.v[e], nv.v[e], dn.v[e], up.v[e], Au.v[e], idg.v[e]);
            }
        }
        if constexpr (ZG) {
            const TV<T, V> bb = ldv<T, V>(&bring[(k + 3) % SB][r][V * lane]);
            out.template pointvp<V>(q, uc, Au, idg, &bb, a0, a1);
        } else if constexpr (NP > 0) {
            out.template pointvp<V>(q, uc, Au, idg, co, a0, a1);
        } else {
            out.template pointv<V>(q, uc, Au, idg, a0, a1);
        }
        if (i0 == 1) {                               // the boundary point i = 0
            const T t0 = thalo[tc][r][0];
            T A0, i00;
            op.bnd(0, j, k, t0, A0, i00);
            if constexpr (ZG) {
                const T x0 = bhalo[(k + 3) % SB][r][0];
                out.pointp(q - 1, t0, A0, i00, &x0, a0, a1);
            } else if constexpr (NP > 0) {
                out.pointp(q - 1, t0, A0, i00, cox, a0, a1);
            } else {
                out.point(q - 1, t0, A0, i00, a0, a1);
            }
        }
        ud = uc;
        uc = uu;
#pragma unroll
        for (int qq = 0; qq < NPA; ++qq) {
            co[qq] = ro[qq];
            cox[qq] = rox[qq];
        }
        q += sm;
    }
}

template <typename T, int V, class Op, bool ZG, class Out, int TYD>
__global__ void __launch_bounds__(TX * TYD)
k_prolong_smooth_lds(int mx, int my, int mz, int mxc, int myc, int kch, Op op, const T* __restrict__ xb, T c,
                     const T* __restrict__ xc, Out out, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    const Tile tl = xcd_tile();
    prolong_smooth_lds<T, V, Op, ZG, Out, TYD>(tl.bx, tl.by, tl.bz, mx, my, mz, mxc, myc, kch, op, xb, c, xc, out,
                                               a0, a1);
    block_partials<Out::NDOT, TYD>(a0, a1, part, tl.id);
}

// ------------------------------------------------------- coarse solve

// x = Ainv b, Ainv = A_0^-1 (n x n, row major, FP64): the coarsest-grid
// direct solve as one dense product.  A wave per row, lanes over columns
// (coalesced), fixed-order wave sums; rows row0, row0 + rstride, ...
// (row0 uniform per wave).
template <typename T>
__device__ __forceinline__ void coarse_rows(int n, const double* __restrict__ Ainv, const T* __restrict__ b,
                                            T* __restrict__ x, int row0, int rstride) {
    const int lane = threadIdx.x & 63;
    for (int r = row0; r < n; r += rstride) {
        const double* __restrict__ A = Ainv + (long)r * n;
        double s = 0.0;
        for (int c = lane; c < n; c += 64) s += A[c] * double(b[c]);
        s = wave_sum(s);
        if (lane == 0) x[r] = T(s);
    }
}

// launch: <<<ceil(n/4), 256>>>
template <typename T>
__global__ void __launch_bounds__(256)
k_coarse(int n, const double* __restrict__ Ainv, const T* __restrict__ b, T* __restrict__ x) {
    coarse_rows<T>(n, Ainv, b, x, blockIdx.x * 4 + (threadIdx.x >> 6), gridDim.x * 4);
}

// ------------------------------------------------------- vector kernels

// y = c x (pointwise), optional dots acc0 += y.x, acc1 += y.y
template <typename T, int DOTS>
__global__ void __launch_bounds__(256)
k_scale(long n, T c, const T* __restrict__ x, T* __restrict__ y, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const T xv = x[q];
        const T v = c * xv;
        y[q] = v;
        if (DOTS) { a0 += double(v) * double(xv); a1 += double(v) * double(v); }
    }
    if (DOTS) {
        __shared__ double sm[2][4];
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        if ((threadIdx.x & 63) == 0) { sm[0][threadIdx.x >> 6] = a0; sm[1][threadIdx.x >> 6] = a1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = sm[0][0] + sm[0][1] + sm[0][2] + sm[0][3];
            part[2 * blockIdx.x + 1] = sm[1][0] + sm[1][1] + sm[1][2] + sm[1][3];
        }
    }
}

// x converted to y (a narrower type); acc0 = acc1 += x.x
template <typename S, typename D>
__global__ void __launch_bounds__(256) k_convert_norm(long n, const S* __restrict__ x, D* __restrict__ y,
                                                     double* __restrict__ part) {
    double a0 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double xv = double(x[q]);
        y[q] = D(x[q]);
        a0 += xv * xv;
    }
    __shared__ double sm[4];
    a0 = wave_sum(a0);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a0;
    __syncthreads();
    if (threadIdx.x == 0) part[2 * blockIdx.x] = part[2 * blockIdx.x + 1] = sm[0] + sm[1] + sm[2] + sm[3];
}

// acc0 += x.y, acc1 += x.x
template <typename T>
__global__ void __launch_bounds__(256)
k_dot2(long n, const T* __restrict__ x, const T* __restrict__ y, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const double xv = double(x[q]);
        a0 += xv * double(y[q]);
        a1 += xv * xv;
    }
    __shared__ double sm[2][4];
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    if ((threadIdx.x & 63) == 0) { sm[0][threadIdx.x >> 6] = a0; sm[1][threadIdx.x >> 6] = a1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0][0] + sm[0][1] + sm[0][2] + sm[0][3];
        part[2 * blockIdx.x + 1] = sm[1][0] + sm[1][1] + sm[1][2] + sm[1][3];
    }
}

// p_new = z + beta p_old and the deferred iterate update x += alpha p_old
// (beta, alpha device scalars; X may be wider than T)
template <typename T, typename X>
__global__ void __launch_bounds__(256)
k_dir_x(long n, const double* __restrict__ beta, const double* __restrict__ alpha, const T* __restrict__ z,
        const T* __restrict__ pold, T* __restrict__ pnew, X* __restrict__ x, int init) {
    const T bt = T(*beta);
    const X al = X(*alpha);
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const T p = pold[q];
        pnew[q] = z[q] + bt * p;
        x[q] = (init ? X(0) : x[q]) + al * X(p);
    }
}

// rn = r - a w (a = S[ialpha]); NORM: acc0 += rn.rn
template <typename T, int NORM>
__global__ void __launch_bounds__(256)
k_r_update(long n, const double* __restrict__ S, int ialpha, const T* __restrict__ r, const T* __restrict__ w,
           T* __restrict__ rn, double* __restrict__ part) {
    const T a = T(S[ialpha]);
    double acc = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const T v = r[q] - a * w[q];
        rn[q] = v;
        if (NORM) acc += double(v) * double(v);
    }
    if (NORM) {
        __shared__ double sm[4];
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
            part[2 * blockIdx.x + 1] = 0.0;
        }
    }
}

// x += alpha_0 p_0 + alpha_1 p_1 + ... + alpha_{M-1} p_{M-1}, one term at a
// time in that order (the rounding of per-iteration updates): deferred
// iterate updates, four points per thread (16-byte accesses), M = f.m;
// init: x taken as 0, not read
template <int M, typename P>
__global__ void __launch_bounds__(256) k_fold_x(long n, double* __restrict__ x, FoldListT<P> f, int init) {
    double al[M];
#pragma unroll
    for (int j = 0; j < M; ++j) al[j] = *f.a[j];
    const long n4 = n / 4;
    for (long t = blockIdx.x * 256L + threadIdx.x; t < n4; t += (long)gridDim.x * 256) {
        const long q = 4 * t;
        TV<double, 2> u0, u1;
        if (init) {
            u0.v[0] = u0.v[1] = u1.v[0] = u1.v[1] = 0.0;
        } else {
            u0 = ldv<double, 2>(x + q);
            u1 = ldv<double, 2>(x + q + 2);
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
            P pp[4];
            if constexpr (sizeof(P) == 4) {
                const TV<P, 4> v = ldv<P, 4>(f.p[j] + q);
                pp[0] = v.v[0]; pp[1] = v.v[1]; pp[2] = v.v[2]; pp[3] = v.v[3];
            } else {
                const TV<P, 2> v0 = ldv<P, 2>(f.p[j] + q), v1 = ldv<P, 2>(f.p[j] + q + 2);
                pp[0] = v0.v[0]; pp[1] = v0.v[1]; pp[2] = v1.v[0]; pp[3] = v1.v[1];
            }
            u0.v[0] += al[j] * double(pp[0]);
            u0.v[1] += al[j] * double(pp[1]);
            u1.v[0] += al[j] * double(pp[2]);
            u1.v[1] += al[j] * double(pp[3]);
        }
        stv<double, 2>(x + q, u0);
        stv<double, 2>(x + q + 2, u1);
    }
    for (long q = 4 * n4 + blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        double u = init ? 0.0 : x[q];
#pragma unroll
        for (int j = 0; j < M; ++j) u += al[j] * double(f.p[j][q]);
        x[q] = u;
    }
}

// *dst = *src (a device scalar saved before the slot is overwritten)
__global__ void __launch_bounds__(64) k_copy_scalar(const double* __restrict__ src, double* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
}

// x += (double) c   (fold an FP32 correction into an FP64 iterate)
__global__ void __launch_bounds__(256)
k_axpy_mixed(long n, double* __restrict__ x, const float* __restrict__ c) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) x[q] += double(c[q]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(long n, T v, T* __restrict__ x) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) x[q] = v;
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) k_convert(long n, const S* __restrict__ x, D* __restrict__ y) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256) y[q] = D(x[q]);
}

// deterministic start vector of the Chebyshev eigen-estimate (splitmix64)
__device__ __forceinline__ double noisy_at(long index) {
    uint64_t z = (uint64_t)index + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256) k_noisy(long n, double* __restrict__ x) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256)
        x[q] = noisy_at(q);
}

// ------------------------------------------------------- finalisation

// Sum nparts pairs in a fixed order, then update the scalar block S.
// Scalar slots:
enum {
    S_RZ = 0,      // beta = (z, r)
    S_RZOLD,       // betaold
    S_PW,          // dpi = (p, A p)
    S_ALPHA,       // beta / dpi
    S_BCOEF,       // beta / betaold (direction update)
    S_DP,          // norm of the convergence test
    S_ZZ,          // (z, z)
    S_RR,          // (r, r)
    S_TMP0, S_TMP1,
    S_ONE,         // constant 1 (set at set-up)
    S_ZERO,        // constant 0
    S_COUNT = 16
};
enum { FIN_ALPHA = 0, FIN_BETA_PRECOND, FIN_BETA_NATURAL, FIN_BETA_UNPRECOND, FIN_RR, FIN_RAW };

// scalar-block update from the two reduced sums A, B
__device__ __forceinline__ void finalize_scalars(double* __restrict__ S, int mode, int first, double A,
                                                 double B) {
    switch (mode) {
    case FIN_ALPHA:            // A = (p, A p)
        S[S_PW] = A;
        S[S_RZOLD] = S[S_RZ];
        S[S_ALPHA] = S[S_RZ] / A;
        break;
    case FIN_BETA_PRECOND:     // A = (z, r), B = (z, z)
    case FIN_BETA_NATURAL:
    case FIN_BETA_UNPRECOND:
        S[S_RZ] = A;
        S[S_ZZ] = B;
        if (mode == FIN_BETA_PRECOND) S[S_DP] = sqrt(B);
        else if (mode == FIN_BETA_NATURAL) S[S_DP] = sqrt(fabs(A));
        else S[S_DP] = sqrt(S[S_RR]);
        S[S_BCOEF] = first ? 0.0 : A / S[S_RZOLD];
        break;
    case FIN_RR:               // A = (r, r)
        S[S_RR] = A;
        break;
    case FIN_RAW:              // raw sums into TMP slots
        S[S_TMP0] = A;
        S[S_TMP1] = B;
        break;
    }
}

// Fixed-order sum of nparts (A, B) pairs by one block of FT threads: each
// thread adds a strided subset, then waves and the block reduce in a fixed
// pattern.  The result is on thread 0.
constexpr int FT = 1024;

__device__ __forceinline__ void sum_parts(const double* __restrict__ part, long nparts, double& A, double& B) {
    __shared__ double s0[FT / 64], s1[FT / 64];
    double a = 0.0, b = 0.0;
    for (long q = threadIdx.x; q < nparts; q += FT) {
        const T2<double> v = ld2(part + 2 * q);
        a += v.a;
        b += v.b;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    A = 0.0;
    B = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < FT / 64; ++w) {
            A += s0[w];
            B += s1[w];
        }
}

// launch: <<<1, FT>>>
__global__ void __launch_bounds__(FT)
k_finalize(long nparts, const double* __restrict__ part, double* __restrict__ S, int mode,
           int first) {
    double A, B;
    sum_parts(part, nparts, A, B);
    if (threadIdx.x == 0) finalize_scalars(S, mode, first, A, B);
}


// ------------------------------------------------------ SOR smoothing
// PETSc's MatSOR with omega = 1 and a zero initial guess, symmetric sweep
// (the PCSOR default): forward x_q = (b_q - sum_{j<q} a_qj x_j) / a_qq, then
// backward x_q = (b_q - sum_{j!=q} a_qj x_j) / a_qq with the newest values,
// in lexicographic (x fastest) order.  A vertex depends only on neighbours
// earlier in that order, so every vertex of the hyperplane
// wi*i + wj*j + wk*k = p (weights making every earlier neighbour's plane
// smaller: 1,1,1 for the 3/5/7-point operator, 1,2,4 for 27 points) is
// updated in parallel, plane after plane, in exactly the serial arithmetic.
// Local sweeps of nblk ranks (blocks of the slowest dimension starting at
// blk[0..nblk-1]) drop the couplings between blocks, as PETSc's
// SOR_LOCAL_SYMMETRIC_SWEEP does on several processes.
struct SorOp {
    int dim, mx, my, mz;
    long n;
    double d, db, sx, sy, sz;     // constant operator (g == nullptr)
    const double* g;              // 27-point coefficients (Galerkin or VI levels)
    int nblk;
    int blk[17];
};

__device__ __forceinline__ void sor_block(const SorOp& o, int s, int& lo, int& hi) {
    lo = 0;
    hi = o.dim == 3 ? o.mz : (o.dim == 2 ? o.my : o.mx);
    for (int q = 1; q < o.nblk; ++q) {
        if (o.blk[q] <= s) lo = o.blk[q];
        else { hi = o.blk[q]; break; }
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_sor_plane(SorOp o, int p, int backward, const T* __restrict__ b, T* __restrict__ x) {
    const long t = blockIdx.x * 256L + threadIdx.x;
    const int wj = o.g ? 2 : 1, wk = o.g ? 4 : 1;
    int i, j, k;
    if (o.dim == 3) {
        if (t >= (long)o.my * o.mz) return;
        j = (int)(t % o.my);
        k = (int)(t / o.my);
    } else if (o.dim == 2) {
        if (t >= o.my) return;
        j = (int)t;
        k = 0;
    } else {
        if (t > 0) return;
        j = 0;
        k = 0;
    }
    i = p - wj * j - wk * k;
    if (i < 0 || i >= o.mx) return;
    const int sdim = o.dim == 3 ? k : (o.dim == 2 ? j : i);
    int lo, hi;
    sor_block(o, sdim, lo, hi);
    const long sy = o.mx, sz = (long)o.mx * o.my;
    const long q = ((long)k * o.my + j) * o.mx + i;
    double s = double(b[q]), dq;
    if (o.g) {
        dq = o.g[13 * o.n + q];
        for (int c = 0; c < 27; ++c) {
            if (c == 13 || (!backward && c > 13)) continue;
            const int oz = c / 9 - 1, oy = (c / 3) % 3 - 1, ox = c % 3 - 1;
            const int ii = i + ox, jj = j + oy, kk = k + oz;
            if (ii < 0 || ii >= o.mx || jj < 0 || jj >= o.my || kk < 0 || kk >= o.mz) continue;
            const int sn = o.dim == 3 ? kk : (o.dim == 2 ? jj : ii);
            if (sn < lo || sn >= hi) continue;
            const double a = o.g[(long)c * o.n + q];
            if (a != 0.0) s -= a * double(x[q + oz * sz + oy * sy + ox]);
        }
        x[q] = dq != 0.0 ? T(s * (1.0 / dq)) : T(0);   // a null row stays zero
        return;
    }
    const bool bnd = i == 0 || i == o.mx - 1 || (o.dim >= 2 && (j == 0 || j == o.my - 1)) ||
                     (o.dim == 3 && (k == 0 || k == o.mz - 1));
    if (bnd) {
        x[q] = T(s * (1.0 / o.db));
        return;
    }
    // interior neighbours only, in column order: k-1, j-1, i-1 | i+1, j+1, k+1
    auto in = [&](int ii, int jj, int kk) {
        const bool nb = ii == 0 || ii == o.mx - 1 || (o.dim >= 2 && (jj == 0 || jj == o.my - 1)) ||
                        (o.dim == 3 && (kk == 0 || kk == o.mz - 1));
        const int sn = o.dim == 3 ? kk : (o.dim == 2 ? jj : ii);
        return !nb && sn >= lo && sn < hi;
    };
    if (o.dim == 3 && in(i, j, k - 1)) s += o.sz * double(x[q - sz]);
    if (o.dim >= 2 && in(i, j - 1, k)) s += o.sy * double(x[q - sy]);
    if (in(i - 1, j, k)) s += o.sx * double(x[q - 1]);
    if (backward) {
        if (in(i + 1, j, k)) s += o.sx * double(x[q + 1]);
        if (o.dim >= 2 && in(i, j + 1, k)) s += o.sy * double(x[q + sy]);
        if (o.dim == 3 && in(i, j, k + 1)) s += o.sz * double(x[q + sz]);
    }
    x[q] = T(s * (1.0 / o.d));
}

// ILU(0) of the constant 3/5/7-point operator (PETSc's PCILU / PCICC with
// no fill): the factor keeps A's pattern and A's off-diagonal entries; only
// the pivots u change (computed on the host in lexicographic order).  The
// solves run on the same hyperplanes as the SOR sweeps: forward
// y_q = r_q - sum_{lower m} (a_qm / u_m) y_m, backward
// z_q = (y_q - sum_{upper m} a_qm z_m) / u_q, each sum in column order.
template <typename T>
__global__ void __launch_bounds__(256)
k_ilu_plane(SorOp o, const double* __restrict__ u, int p, int backward, const T* __restrict__ r,
            T* __restrict__ z) {
    const long t = blockIdx.x * 256L + threadIdx.x;
    int i, j, k;
    if (o.dim == 3) {
        if (t >= (long)o.my * o.mz) return;
        j = (int)(t % o.my);
        k = (int)(t / o.my);
    } else if (o.dim == 2) {
        if (t >= o.my) return;
        j = (int)t;
        k = 0;
    } else {
        if (t > 0) return;
        j = 0;
        k = 0;
    }
    i = p - j - k;
    if (i < 0 || i >= o.mx) return;
    const long sy = o.mx, sz = (long)o.mx * o.my;
    const long q = ((long)k * o.my + j) * o.mx + i;
    auto bnd = [&](int ii, int jj, int kk) {
        return ii == 0 || ii == o.mx - 1 || (o.dim >= 2 && (jj == 0 || jj == o.my - 1)) ||
               (o.dim == 3 && (kk == 0 || kk == o.mz - 1));
    };
    const bool bq = bnd(i, j, k);
    if (!backward) {
        double s = double(r[q]);
        if (!bq) {   // a = -s toward interior neighbours
            if (o.dim == 3 && !bnd(i, j, k - 1)) s -= ((-o.sz) / u[q - sz]) * double(z[q - sz]);
            if (o.dim >= 2 && !bnd(i, j - 1, k)) s -= ((-o.sy) / u[q - sy]) * double(z[q - sy]);
            if (!bnd(i - 1, j, k)) s -= ((-o.sx) / u[q - 1]) * double(z[q - 1]);
        }
        z[q] = T(s);
        return;
    }
    double s = double(z[q]);
    if (!bq) {
        if (!bnd(i + 1, j, k)) s -= (-o.sx) * double(z[q + 1]);
        if (o.dim >= 2 && !bnd(i, j + 1, k)) s -= (-o.sy) * double(z[q + sy]);
        if (o.dim == 3 && !bnd(i, j, k + 1)) s -= (-o.sz) * double(z[q + sz]);
    }
    z[q] = T(s / u[q]);
}

// ILU(0) of an operator in 27-point form whose nonzeros lie on the 7-point
// pattern (the null-row operator of a reduced VI step): no fill, so the
// factor is A's pattern with new pivots.  Pivots, plane by plane (i+j+k):
// u_q = a_qq - sum_{lower m} (a_qm / u_m) a_mq in column order; a null row
// keeps u_q = 0 and its solves give 0.
__device__ __forceinline__ bool ilu27_site(const SorOp& o, long t, int p, int& i, int& j, int& k) {
    if (o.dim == 3) {
        if (t >= (long)o.my * o.mz) return false;
        j = (int)(t % o.my);
        k = (int)(t / o.my);
    } else if (o.dim == 2) {
        if (t >= o.my) return false;
        j = (int)t;
        k = 0;
    } else {
        if (t > 0) return false;
        j = 0;
        k = 0;
    }
    i = p - j - k;
    return i >= 0 && i < o.mx;
}

__global__ void __launch_bounds__(256) k_ilu27_pivot_plane(SorOp o, double* __restrict__ u, int p) {
    int i, j, k;
    if (!ilu27_site(o, blockIdx.x * 256L + threadIdx.x, p, i, j, k)) return;
    const long sy = o.mx, sz = (long)o.mx * o.my;
    const long q = ((long)k * o.my + j) * o.mx + i;
    double v = o.g[13 * o.n + q];
    if (v != 0.0) {
        // lower neighbours in column order: k-1 (c 4), j-1 (c 10), i-1 (c 12);
        // a_mq is row m's coefficient toward q (c 22, 16, 14)
        if (k > 0 && o.g[4 * o.n + q] != 0.0) v -= (o.g[4 * o.n + q] / u[q - sz]) * o.g[22 * o.n + q - sz];
        if (j > 0 && o.g[10 * o.n + q] != 0.0) v -= (o.g[10 * o.n + q] / u[q - sy]) * o.g[16 * o.n + q - sy];
        if (i > 0 && o.g[12 * o.n + q] != 0.0) v -= (o.g[12 * o.n + q] / u[q - 1]) * o.g[14 * o.n + q - 1];
    }
    u[q] = v;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_ilu27_plane(SorOp o, const double* __restrict__ u, int p, int backward, const T* __restrict__ r,
              T* __restrict__ z) {
    int i, j, k;
    if (!ilu27_site(o, blockIdx.x * 256L + threadIdx.x, p, i, j, k)) return;
    const long sy = o.mx, sz = (long)o.mx * o.my;
    const long q = ((long)k * o.my + j) * o.mx + i;
    const double uq = u[q];
    if (uq == 0.0) { z[q] = T(0); return; }
    if (!backward) {
        double s = double(r[q]);
        if (k > 0 && o.g[4 * o.n + q] != 0.0) s -= (o.g[4 * o.n + q] / u[q - sz]) * double(z[q - sz]);
        if (j > 0 && o.g[10 * o.n + q] != 0.0) s -= (o.g[10 * o.n + q] / u[q - sy]) * double(z[q - sy]);
        if (i > 0 && o.g[12 * o.n + q] != 0.0) s -= (o.g[12 * o.n + q] / u[q - 1]) * double(z[q - 1]);
        z[q] = T(s);
        return;
    }
    double s = double(z[q]);
    if (i + 1 < o.mx && o.g[14 * o.n + q] != 0.0) s -= o.g[14 * o.n + q] * double(z[q + 1]);
    if (j + 1 < o.my && o.g[16 * o.n + q] != 0.0) s -= o.g[16 * o.n + q] * double(z[q + sy]);
    if (k + 1 < o.mz && o.g[22 * o.n + q] != 0.0) s -= o.g[22 * o.n + q] * double(z[q + sz]);
    z[q] = T(s / uq);
}

// One directional sweep of PETSc's SOR (omega 1, zero initial guess) on a
// 2D 27-plane level (planes 9..17), as strips of 64 rows, one wave each.
// Lane r owns row j of its strip and at step t takes the point c = t - 2r
// along its row (sweep order), so the row below (forward) or above
// (backward) is one lane over and two steps ahead: its three values reach
// lane r by shuffles of each lane's last three results.  Lane 0 takes that
// row from the previous strip through global memory, after that strip's
// last lane has published (release) how many points of its row are done;
// a strip starts SOR9_SLACK points behind, so in steady state it rarely
// polls.  The arithmetic is k_sor_plane's: forward x = (b - sum over
// c = 9..12 of a_c x_c) / a_13; backward the same over c = 9..12, 14..17,
// the lower neighbours holding forward values; a null row stays zero.
// flag: per strip, base + points done (bases grow from sweep to sweep); a
// wait longer than spin_limit polls sets *err and ends the waiting (no
// hang), and a launch that finds *err already set waits for nothing: after
// a stall every later sweep runs through at once until the host, which
// checks *err after each linear solve, stops the solve.  need_extra is added
// to every wait's target (0; the host's stall test sets a target no strip
// reaches, so that every wait gives up after spin_limit polls).
constexpr int SOR9_SLACK = 64;

// lane i receives lane i-1's value (DPP wave_shr:1; lane 0 keeps its own)
__device__ __forceinline__ double wave_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const int rlo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}

__device__ __forceinline__ double ld_agent(const double* p) {
    const unsigned long long v =
        __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

// the per-point record of k_sor9_strips: the 2D planes c = 9..17 with the
// diagonal replaced by its reciprocal (0 for a null row), and a pad
__global__ void __launch_bounds__(256)
k_sor9_pack(long n, const double* __restrict__ g, double* __restrict__ pk) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        double* o = pk + q * 10;
#pragma unroll
        for (int c = 0; c < 9; ++c) o[c] = g[(long)(9 + c) * n + q];
        const double dq = g[13 * n + q];
        o[4] = dq != 0.0 ? 1.0 / dq : 0.0;
        o[9] = dq;
    }
}

// the operands of one step of k_sor9_strips, loaded SOR9_AHEAD steps early
struct Sor9Ops {
    double b, a[10], x[4];
};
constexpr int SOR9_AHEAD = 4;

__global__ void __launch_bounds__(64)
k_sor9_strips(int mx, int my, const double* __restrict__ pk, const double* __restrict__ b, double* x,
              unsigned long long* flag, unsigned long long base, int backward, int* err, long spin_limit,
              unsigned long long need_extra) {
    const int s = blockIdx.x, r = threadIdx.x;
    const int row0 = 64 * s;
    const int rows = min(64, my - row0);
    const int j = backward ? my - 1 - (row0 + r) : row0 + r;
    const bool live = r < rows;
    const bool hs = j > 0, hn = j + 1 < my;
    const int nsteps = mx + 2 * (rows - 1);
    // the previous strip's last row, copied into an LDS ring by the wave in
    // coalesced batches once that strip has published it (lane 0's row
    // neighbour); positions are along the sweep: p -> column p or mx-1-p
    __shared__ double ring[256];
    const int jn = backward ? my - row0 : row0 - 1;   // lane 0's neighbour row (s > 0), the same for all lanes
    int copied = 0;                                    // positions in the ring (wave-uniform)
    bool dead = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    double h1 = 0.0, h2 = 0.0, h3 = 0.0, xo = 0.0;
    Sor9Ops pf[SOR9_AHEAD];   // a ring of prefetched operands, indexed by step mod SOR9_AHEAD
    auto fetch = [&](int t, Sor9Ops& o) {
        const int c = t - 2 * r;
        if (!(live && c >= 0 && c < mx)) return;
        const int i = backward ? mx - 1 - c : c;
        const long q = (long)j * mx + i;
        o.b = b[q];
        const double* rec = pk + q * 10;
#pragma unroll
        for (int k = 0; k < 10; ++k) o.a[k] = rec[k];
        if (backward) {   // forward values below and to the left
            o.x[0] = (hs && i > 0) ? x[q - mx - 1] : 0.0;
            o.x[1] = hs ? x[q - mx] : 0.0;
            o.x[2] = (hs && i + 1 < mx) ? x[q - mx + 1] : 0.0;
            o.x[3] = i > 0 ? x[q - 1] : 0.0;
        }
    };
    auto step = [&](int t, Sor9Ops& cur) {
        // lane 0 is at position t: it needs positions t-1..t+1 of the neighbour row
        if (s > 0 && t < mx && copied < min(t + 2, mx)) {   // uniform
            unsigned long long kn = 0;
            if (r == 0) {
                const unsigned long long need =
                    base + (unsigned long long)min(t == 0 ? SOR9_SLACK : t + 2, mx) + need_extra;
                long spins = 0;
                // relaxed polls (an acquire per poll would invalidate the
                // cache every time), one acquire fence once the count is there
                while (!dead && (kn = __hip_atomic_load(flag + s - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
                    if (++spins > spin_limit) {
                        atomicExch(err, 1);
                        dead = true;
                    } else if ((spins & 63) == 0 &&
                               __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                        dead = true;   // another workgroup gave up: stop waiting too
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (dead) kn = base + (unsigned long long)mx;
            }
            kn = __shfl(kn, 0, 64);
            const int avail = (int)min((unsigned long long)mx, kn - base);
            const int upto = min(avail, t + 130);
            for (int p = copied + r; p < upto; p += 64) {
                const int col = backward ? mx - 1 - p : p;
                ring[p & 255] = ld_agent(x + (long)jn * mx + col);
            }
            __syncthreads();
            copied = upto;
        }
        const double u1 = wave_shr1(h1), u2 = wave_shr1(h2), u3 = wave_shr1(h3);
        const int c = t - 2 * r;
        double res = 0.0;
        if (live && c >= 0 && c < mx) {
            const double* a = cur.a;
            const int i = backward ? mx - 1 - c : c;
            const long q = (long)j * mx + i;
            const bool hv = backward ? hn : hs;   // the row toward the sweep's origin
            double v_im1 = 0.0, v_i = 0.0, v_ip1 = 0.0;
            if (hv) {
                if (r > 0) {
                    if (!backward) { v_im1 = u3; v_i = u2; v_ip1 = u1; }
                    else { v_im1 = u1; v_i = u2; v_ip1 = u3; }
                } else {   // positions c-1, c, c+1 of the ring
                    const double pm = c > 0 ? ring[(c - 1) & 255] : 0.0;
                    const double pp = c + 1 < mx ? ring[(c + 1) & 255] : 0.0;
                    v_i = ring[c & 255];
                    if (!backward) { v_im1 = pm; v_ip1 = pp; }
                    else { v_im1 = pp; v_ip1 = pm; }
                }
            }
            const bool hw = i > 0, he = i + 1 < mx;
            double sum = cur.b;
            if (!backward) {
                if (hs && hw && a[0] != 0.0) sum -= a[0] * v_im1;
                if (hs && a[1] != 0.0) sum -= a[1] * v_i;
                if (hs && he && a[2] != 0.0) sum -= a[2] * v_ip1;
                if (hw && a[3] != 0.0) sum -= a[3] * xo;
            } else {
                if (hs && hw && a[0] != 0.0) sum -= a[0] * cur.x[0];
                if (hs && a[1] != 0.0) sum -= a[1] * cur.x[1];
                if (hs && he && a[2] != 0.0) sum -= a[2] * cur.x[2];
                if (hw && a[3] != 0.0) sum -= a[3] * cur.x[3];
                if (he && a[5] != 0.0) sum -= a[5] * xo;
                if (hn && hw && a[6] != 0.0) sum -= a[6] * v_im1;
                if (hn && a[7] != 0.0) sum -= a[7] * v_i;
                if (hn && he && a[8] != 0.0) sum -= a[8] * v_ip1;
            }
            res = a[9] != 0.0 ? sum * a[4] : 0.0;
            x[q] = res;
            xo = res;
            // a full strip publishes its last row's progress every 32 points
            // (the consumer reads this lane's row only: its own release store)
            if (r == 63 && ((c & 31) == 31 || c == mx - 1))
                __hip_atomic_store(flag + s, base + (unsigned long long)(c + 1), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        fetch(t + SOR9_AHEAD, cur);   // this slot now holds step t + SOR9_AHEAD
        h3 = h2;
        h2 = h1;
        h1 = res;
    };
#pragma unroll
    for (int k = 0; k < SOR9_AHEAD; ++k) fetch(k, pf[k]);
    for (int t0 = 0; t0 < nsteps; t0 += SOR9_AHEAD) {
#pragma unroll
        for (int k = 0; k < SOR9_AHEAD; ++k)
            if (t0 + k < nsteps) step(t0 + k, pf[k]);
    }
}

// One direction of a symmetric SOR sweep (omega 1, zero initial guess) of a
// 3D constant-coefficient 7-point level in lexicographic order, pipelined:
// a 64-lane workgroup owns a strip of 64 rows of one plane (rows in sweep
// order: j = 64 s + r forward, my - 1 - 64 s - r backward), lane r its row,
// and marches along the row one point per step, lane r one step behind lane
// r-1, so the neighbour toward the sweep's origin in the same plane is lane
// r-1's result of the previous step (a DPP shift) and the one in the same
// row the lane's own.  Workgroup g = nstrips k' + s (k' = k forward, mz-1-k
// backward) needs the strip before it in its plane (lane 0's neighbour row)
// and the same strip in the plane before it (every lane's neighbour plane):
// both have lower indices.  Steps run in batches of SOR7_BATCH: before a
// batch the first lane waits (relaxed polls, then an agent-scope acquire
// fence) until those two have
// published enough steps, then every lane loads the batch's operands (b,
// the neighbour plane's new values, lane 0's neighbour row; backward also
// the three lower neighbours, which still hold forward values: the points
// after a point in the backward order run later), computes the batch, fences
// its stores (one agent-scope release fence), and the last lane publishes
// the steps done.
// Forward: x = (b + sz x_{k-1} + sy x_{j-1} + sx x_{i-1}) / d over interior
// neighbours, in k_sor_plane's order; backward: (b + sz x_{k-1} + sy
// x_{j-1} + sx x_{i-1} + sx x_{i+1} + sy x_{j+1} + sz x_{k+1}) / d; a
// boundary row b / db.  flag: per workgroup, base + steps done (bases grow
// from sweep to sweep); a wait longer than spin_limit polls sets *err and
// gives up, and a launch that finds *err set (at entry, or every 64 polls
// of a wait) waits for nothing (the host checks *err after the solve); need_extra is added to every wait's target
// (0; the host's stall test sets one no workgroup reaches).  Grid nstrips *
// mz, block 64.
constexpr int SOR7_BATCH = 8;
template <typename T>
__global__ void __launch_bounds__(64)
k_sor7_strips(int mx, int my, int mz, double d, double db, double sx, double sy, double sz,
              const T* __restrict__ b, T* x, unsigned long long* flag, unsigned long long base, int backward,
              int* err, long spin_limit, unsigned long long need_extra) {
    static_assert(sizeof(T) == 8, "FP64 sweeps (the neighbours' new values are read as 8-byte atomics)");
    constexpr int B = SOR7_BATCH;
    const int nstrips = (my + 63) / 64;
    const int g = blockIdx.x, r = threadIdx.x;
    const int s = g % nstrips, kk = g / nstrips;        // strip and plane in sweep order
    const int k = backward ? mz - 1 - kk : kk;
    const int jr = 64 * s + r;                          // row in sweep order
    const int j = backward ? my - 1 - jr : jr;
    const bool live = jr < my;
    const int jn = backward ? j + 1 : j - 1;            // neighbour row toward the origin
    const int kn = backward ? k + 1 : k - 1;            // neighbour plane toward the origin
    const long dy = backward ? (long)mx : -(long)mx, dz = backward ? (long)mx * my : -((long)mx * my);
    const long sy_ = mx, sz_ = (long)mx * my;
    const bool bndjk = j == 0 || j == my - 1 || k == 0 || k == mz - 1;
    auto inner = [&](int ii, int jj, int kq) {
        return ii > 0 && ii < mx - 1 && jj > 0 && jj < my - 1 && kq > 0 && kq < mz - 1;
    };
    const int nsteps = mx + 63;
    bool dead = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    const unsigned long long* fs = s > 0 ? flag + (g - 1) : nullptr;          // previous strip, same plane
    const unsigned long long* fk = kk > 0 ? flag + (g - nstrips) : nullptr;   // same strip, previous plane
    double prev = 0.0;                                  // the lane's result of the previous step
    for (int t0 = 0; t0 < nsteps; t0 += B) {
        const int t1 = min(t0 + B, nsteps);
        // the batch needs the previous strip through its step t1 - 1 + 64 (its
        // last lane at this strip's lane-0 columns) and the previous plane's
        // strip through its step t1 - 1
        if (r == 0) {
            const unsigned long long need_s = base + (unsigned long long)min(t1 + 64, nsteps) + need_extra;
            const unsigned long long need_k = base + (unsigned long long)t1 + need_extra;
            long spins = 0;
            // relaxed polls, then one acquire fence (an acquire per poll would
            // invalidate the cache every time)
            while (!dead && ((fs && __hip_atomic_load(fs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need_s) ||
                             (fk && __hip_atomic_load(fk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need_k))) {
                if (++spins > spin_limit) {
                    atomicExch(err, 1);
                    dead = true;
                } else if ((spins & 63) == 0 &&
                           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    dead = true;   // another workgroup gave up: stop waiting too
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        // the batch's operands
        double bq[B], vk[B], vj[B], lk[B], lj[B], li[B];
#pragma unroll
        for (int e = 0; e < B; ++e) {
            bq[e] = vk[e] = vj[e] = lk[e] = lj[e] = li[e] = 0.0;
            const int c = t0 + e - r;
            if (!(t0 + e < t1 && live && c >= 0 && c < mx)) continue;
            const int i = backward ? mx - 1 - c : c;
            const long q = ((long)k * my + j) * mx + i;
            bq[e] = double(b[q]);
            if (bndjk || i == 0 || i == mx - 1) continue;
            if (inner(i, j, kn)) vk[e] = ld_agent((const double*)(x + q + dz));
            if (r == 0 && inner(i, jn, k)) vj[e] = ld_agent((const double*)(x + q + dy));
            if (backward) {
                if (inner(i, j, k - 1)) lk[e] = double(x[q - sz_]);
                if (inner(i, j - 1, k)) lj[e] = double(x[q - sy_]);
                if (inner(i - 1, j, k)) li[e] = double(x[q - 1]);
            }
        }
#pragma unroll
        for (int e = 0; e < B; ++e) {
            const int t = t0 + e;
            const double nbv = wave_shr1(prev);         // lane r-1's result of step t-1: (i, jn)
            const int c = t - r;
            double res = prev;
            if (t < t1 && live && c >= 0 && c < mx) {
                const int i = backward ? mx - 1 - c : c;
                const int in_ = backward ? i + 1 : i - 1;
                const long q = ((long)k * my + j) * mx + i;
                double sum = bq[e];
                if (bndjk || i == 0 || i == mx - 1) {
                    res = sum * (1.0 / db);
                } else {
                    const bool hk = inner(i, j, kn), hj = inner(i, jn, k), hi = inner(in_, j, k);
                    const double xj = r > 0 ? nbv : vj[e];
                    if (!backward) {
                        if (hk) sum += sz * vk[e];
                        if (hj) sum += sy * xj;
                        if (hi) sum += sx * prev;
                    } else {
                        if (inner(i, j, k - 1)) sum += sz * lk[e];
                        if (inner(i, j - 1, k)) sum += sy * lj[e];
                        if (inner(i - 1, j, k)) sum += sx * li[e];
                        if (hi) sum += sx * prev;
                        if (hj) sum += sy * xj;
                        if (hk) sum += sz * vk[e];
                    }
                    res = sum * (1.0 / d);
                }
                x[q] = T(res);
            }
            prev = res;
        }
        // every lane's stores of the batch before the published count: one
        // release fence for the wave, then the count
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (r == 63)
            __hip_atomic_store(flag + g, base + (unsigned long long)t1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The dense inverse of a band LU factor (unit lower L and upper U in band
// storage, row i at B[i * w + (col - i + bw)], w = 2 bw + 1): thread c
// solves L U z = e_c by band substitution and writes column c of the
// row-major inverse, which is also its workspace (coalesced across c).
__global__ void __launch_bounds__(256)
k_band_inverse(int n, int bw, const double* __restrict__ B, double* __restrict__ Ainv) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int w = 2 * bw + 1;
    for (int i = 0; i < n; ++i) {
        double sum = (i == c) ? 1.0 : 0.0;
        for (int j = max(0, i - bw); j < i; ++j) sum -= B[(long)i * w + (j - i + bw)] * Ainv[(long)j * n + c];
        Ainv[(long)i * n + c] = sum;
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = Ainv[(long)i * n + c];
        for (int j = i + 1; j <= min(i + bw, n - 1); ++j) sum -= B[(long)i * w + (j - i + bw)] * Ainv[(long)j * n + c];
        Ainv[(long)i * n + c] = sum / B[(long)i * w + bw];
    }
}

// ILU(0) of a 2D operator on the full 9-point box pattern (the DMDA box
// stencil's matrix), natural row order, the IKJ elimination of PETSc's
// MatLUFactorNumeric over the ILU(0) pattern.  Row q = (i, j) eliminates its
// lower neighbours SW, S, SE, W in column order; each updates the entries of
// row q that its own upper part reaches.  Row q depends on rows on the
// hyperplanes i + 2j - 1, -2, -3, so a plane p = i + 2j is one launch.  The
// factor f holds 9 planes: L multipliers SW, S, SE, W (0..3), then U's
// diagonal, E, NW, N, NE (4..8).  The operator is g27's 2D planes 9..17.
__global__ void __launch_bounds__(256)
k_ilu9_factor_plane(int mx, int my, const double* __restrict__ g, double* __restrict__ f, int p) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= my) return;
    const int i = p - 2 * j;
    if (i < 0 || i >= mx) return;
    const long n = (long)mx * my, q = (long)j * mx + i;
    auto A = [&](int c) { return g[(long)c * n + q]; };
    auto F = [&](int c, long r) { return f[(long)c * n + r]; };
    const bool hw = i > 0, he = i + 1 < mx, hs = j > 0, hn = j + 1 < my;
    double aS = hs ? A(10) : 0.0, aSE = (hs && he) ? A(11) : 0.0, aW = hw ? A(12) : 0.0, aC = A(13);
    double aE = he ? A(14) : 0.0, aNW = (hn && hw) ? A(15) : 0.0, aN = hn ? A(16) : 0.0;
    const double aNE = (hn && he) ? A(17) : 0.0;
    double lSW = 0.0, lS = 0.0, lSE = 0.0, lW = 0.0;
    if (hs && hw) {   // SW: its E is S, its N is W, its NE is q
        const long k = q - mx - 1;
        lSW = A(9) / F(4, k);
        aS -= lSW * F(5, k);
        aW -= lSW * F(7, k);
        aC -= lSW * F(8, k);
    }
    if (hs) {         // S: its E is SE, its NW is W, its N is q, its NE is E
        const long k = q - mx;
        lS = aS / F(4, k);
        if (he) aSE -= lS * F(5, k);
        if (hw) aW -= lS * F(6, k);
        aC -= lS * F(7, k);
        if (he) aE -= lS * F(8, k);
    }
    if (hs && he) {   // SE: its NW is q, its N is E
        const long k = q - mx + 1;
        lSE = aSE / F(4, k);
        aC -= lSE * F(6, k);
        aE -= lSE * F(7, k);
    }
    if (hw) {         // W: its E is q, its N is NW, its NE is N
        const long k = q - 1;
        lW = aW / F(4, k);
        aC -= lW * F(5, k);
        if (hn) aNW -= lW * F(7, k);
        if (hn) aN -= lW * F(8, k);
    }
    f[0 * n + q] = lSW; f[1 * n + q] = lS; f[2 * n + q] = lSE; f[3 * n + q] = lW;
    f[4 * n + q] = aC; f[5 * n + q] = aE; f[6 * n + q] = aNW; f[7 * n + q] = aN; f[8 * n + q] = aNE;
}

// z = (LU)^-1 r: forward on planes p ascending, backward descending
template <typename T>
__global__ void __launch_bounds__(256)
k_ilu9_plane(int mx, int my, const double* __restrict__ f, int p, int backward, const T* __restrict__ r,
             T* __restrict__ z) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= my) return;
    const int i = p - 2 * j;
    if (i < 0 || i >= mx) return;
    const long n = (long)mx * my, q = (long)j * mx + i;
    const bool hw = i > 0, he = i + 1 < mx, hs = j > 0, hn = j + 1 < my;
    if (!backward) {
        double s = double(r[q]);
        if (hs && hw) s -= f[q] * double(z[q - mx - 1]);
        if (hs) s -= f[n + q] * double(z[q - mx]);
        if (hs && he) s -= f[2 * n + q] * double(z[q - mx + 1]);
        if (hw) s -= f[3 * n + q] * double(z[q - 1]);
        z[q] = T(s);
        return;
    }
    double s = double(z[q]);
    if (he) s -= f[5 * n + q] * double(z[q + 1]);
    if (hn && hw) s -= f[6 * n + q] * double(z[q + mx - 1]);
    if (hn) s -= f[7 * n + q] * double(z[q + mx]);
    if (hn && he) s -= f[8 * n + q] * double(z[q + mx + 1]);
    z[q] = T(s / f[4 * n + q]);
}

// z = D^-1 r of a constant-operator level (boundary rows db, interior d)
template <typename T>
__global__ void __launch_bounds__(256)
k_jacobi_const(int dim, int mx, int my, int mz, double d, double db, const T* __restrict__ r, T* __restrict__ z) {
    const long n = (long)mx * my * mz;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % mx);
        const long t = q / mx;
        const int j = (int)(t % my), k = (int)(t / my);
        const bool bnd = i == 0 || i == mx - 1 || (dim >= 2 && (j == 0 || j == my - 1)) ||
                         (dim == 3 && (k == 0 || k == mz - 1));
        z[q] = T(double(r[q]) * (1.0 / (bnd ? db : d)));
    }
}

// y = a p + c q + e z (p or q may be null: 0)
template <typename T>
__global__ void __launch_bounds__(256)
k_comb3(long n, double a, const T* __restrict__ p, double c, const T* __restrict__ q, double e,
        const T* __restrict__ z, T* __restrict__ y) {
    for (long m = blockIdx.x * 256L + threadIdx.x; m < n; m += (long)gridDim.x * 256L) {
        double v = e * double(z[m]);
        if (p) v += a * double(p[m]);
        if (q) v += c * double(q[m]);
        y[m] = T(v);
    }
}

// y += a x
template <typename T>
__global__ void __launch_bounds__(256) k_axpy(long n, T a, const T* __restrict__ x, T* __restrict__ y) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) y[q] += a * x[q];
}

// ------------------------------------------------ reduced-space VI Newton
// (PETSc's SNESVINEWTONRSLS over the constant 3/5/7-point operator).  A
// reduced Newton step is solved in full-size vectors: the inactive block
// keeps the operator's rows and their couplings to inactive vertices; an
// active vertex becomes a null row (all coefficients zero).  With a right-
// hand side that is zero on the active set, a Krylov iterate, a Galerkin
// coarse level (P^T A P of a null-row operator has null rows exactly where
// no inactive fine vertex restricts) and a Jacobi-smoothed cycle stay equal
// to the reduced-space ones; the only place a value reaches an active
// vertex is the prolongation onto the finest level, which is masked.

// the grid and the constant operator of a level (d, db: interior and
// boundary diagonal; sx, sy, sz: couplings between interior vertices)
struct ViGrid {
    int dim, mx, my, mz;
    long n;
    double d, db, sx, sy, sz;
};

__device__ __forceinline__ bool vi_bnd(const ViGrid& g, int i, int j, int k) {
    return i == 0 || i == g.mx - 1 || (g.dim >= 2 && (j == 0 || j == g.my - 1)) ||
           (g.dim == 3 && (k == 0 || k == g.mz - 1));
}

__device__ __forceinline__ void vi_ijk(const ViGrid& g, long q, int& i, int& j, int& k) {
    i = (int)(q % g.mx);
    const long r = q / g.mx;
    j = (int)(r % g.my);
    k = (int)(r / g.my);
}

// (A x)_q of the constant operator: boundary rows db x; interior rows
// d x - s * (interior neighbours)
__device__ __forceinline__ double vi_apply(const ViGrid& g, const double* __restrict__ x, long q, int i, int j,
                                           int k) {
    if (vi_bnd(g, i, j, k)) return g.db * x[q];
    const long sy = g.mx, sz = (long)g.mx * g.my;
    double nb = 0.0;
    if (!vi_bnd(g, i - 1, j, k)) nb += g.sx * x[q - 1];
    if (!vi_bnd(g, i + 1, j, k)) nb += g.sx * x[q + 1];
    if (g.dim >= 2) {
        if (!vi_bnd(g, i, j - 1, k)) nb += g.sy * x[q - sy];
        if (!vi_bnd(g, i, j + 1, k)) nb += g.sy * x[q + sy];
    }
    if (g.dim == 3) {
        if (!vi_bnd(g, i, j, k - 1)) nb += g.sz * x[q - sz];
        if (!vi_bnd(g, i, j, k + 1)) nb += g.sz * x[q + sz];
    }
    return g.d * x[q] - nb;
}

// the active-set predicate of SNESVIGetActiveSetIS (bounds: xl, xu or none)
__device__ __forceinline__ bool vi_is_active(double x, double f, double xl, const double* __restrict__ xu, long q) {
    return (x <= xl + 1.0e-8 && f > 0.0) || (xu && x >= xu[q] - 1.0e-8 && f < 0.0);
}

// F = A x + F0 (F0 null: F = A x)
__global__ void __launch_bounds__(256)
k_vi_fun(ViGrid g, const double* __restrict__ x, const double* __restrict__ F0, double* __restrict__ F) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < g.n; q += (long)gridDim.x * 256L) {
        int i, j, k;
        vi_ijk(g, q, i, j, k);
        const double a = vi_apply(g, x, q, i, j, k);
        F[q] = F0 ? a + F0[q] : a;
    }
}

// x = clamp(x0 - lambda y, xl, xu) (SNESVIProjectOntoBounds of a trial point;
// y null: project x0)
__global__ void __launch_bounds__(256)
k_vi_trial(long n, const double* __restrict__ x0, const double* __restrict__ y, double lambda,
           const double* __restrict__ xl, const double* __restrict__ xu, double* __restrict__ x) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        double v = y ? x0[q] - lambda * y[q] : x0[q];
        if (xl && v < xl[q]) v = xl[q];
        else if (xu && v > xu[q]) v = xu[q];
        x[q] = v;
    }
}

// partials: a0 = sum of F^2 over the inactive set (SNESVIComputeInactiveSet-
// Fnorm), a1 = |x|^2.  With act non-null also the active flags and the
// reduced right-hand side Fm (F on the inactive set, 0 on the active set).
__global__ void __launch_bounds__(256)
k_vi_sets(long n, const double* __restrict__ x, const double* __restrict__ F, const double* __restrict__ xl,
          const double* __restrict__ xu, unsigned char* __restrict__ act, double* __restrict__ Fm,
          double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const double xv = x[q], fv = F[q];
        const bool a = vi_is_active(xv, fv, xl[q], xu, q);
        if (!a) a0 += fv * fv;
        a1 += xv * xv;
        if (act) {
            act[q] = a ? 1 : 0;
            Fm[q] = a ? 0.0 : fv;
        }
    }
    block_partials<2, 4>(a0, a1, part, blockIdx.x);
}

// partials: a0 = (x, y), a1 = (y, y)
__global__ void __launch_bounds__(256)
k_vi_dot(long n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const double yv = y[q];
        a0 += x[q] * yv;
        a1 += yv * yv;
    }
    block_partials<2, 4>(a0, a1, part, blockIdx.x);
}

// PCASM's restricted application: every row a subdomain owns (4 longs per
// row: the offset of its inverse row, of its subdomain's vertex list, the
// subdomain size, the row's vertex) gets the row of the subdomain's inverse
// applied to r gathered over the subdomain; rows no subdomain owns keep z's
// value (the caller fills z with 0 first)
__global__ void __launch_bounds__(256)
k_asm_apply(long nrows, const long* __restrict__ rows, const int* __restrict__ idx, const double* __restrict__ inv,
            const double* __restrict__ r, double* __restrict__ z) {
    for (long k = blockIdx.x * 256L + threadIdx.x; k < nrows; k += (long)gridDim.x * 256L) {
        const long* R = rows + 4 * k;
        const double* a = inv + R[0];
        const int* c = idx + R[1];
        const long cnt = R[2];
        double s = 0.0;
        for (long j = 0; j < cnt; ++j) s += a[j] * r[c[j]];
        z[R[3]] = s;
    }
}

// SNESVINEWTONSSLS (semismooth Newton on the Fischer-Burmeister
// reformulation, lower bounds only as the obstacle problem sets them):
// phi(a, b) = a + b - sqrt(a^2 + b^2) and its derivative in a, with PETSc's
// B-subdifferential choice 1 - sqrt(2)/2 where a = b = 0.  No FMA
// contraction: the same IEEE operations as the plain formula.
__device__ __forceinline__ double fb_phi(double a, double b) {
#pragma clang fp contract(off)
    return a + b - sqrt(a * a + b * b);
}
__device__ __forceinline__ double fb_dphi(double a, double b) {
#pragma clang fp contract(off)
    if (fabs(a) <= 1.0e-30 && fabs(b) <= 1.0e-30) return 1.0 - sqrt(2.0) / 2.0;
    return 1.0 - a / sqrt(a * a + b * b);
}

// phi_q = Phi(x_q - l_q, F_q)
__global__ void __launch_bounds__(256)
k_ssls_phi(long n, const double* __restrict__ x, const double* __restrict__ xl, const double* __restrict__ F,
           double* __restrict__ phi) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L)
        phi[q] = fb_phi(x[q] - xl[q], F[q]);
}

// the semismooth Jacobian diag(Db) J + diag(Da) in 27-point form: J the
// constant star operator of g (Dirichlet rows diagonal-only, interior rows
// coupled to interior neighbours), Da_q = Phi_a(x_q - l_q, F_q), Db_q =
// Phi_a(F_q, x_q - l_q); each row scaled by Db_q, then Da_q added to its
// diagonal (MatDiagonalScale, MatDiagonalSet with ADD_VALUES)
__global__ void __launch_bounds__(256)
k_ssls_op27(ViGrid g, const double* __restrict__ x, const double* __restrict__ xl, const double* __restrict__ F,
            double* __restrict__ g27) {
#pragma clang fp contract(off)
    for (long q = blockIdx.x * 256L + threadIdx.x; q < g.n; q += (long)gridDim.x * 256L) {
        int i, j, k;
        vi_ijk(g, q, i, j, k);
        const double a = x[q] - xl[q], f = F[q];
        const double da = fb_dphi(a, f), db = fb_dphi(f, a);
        double c[27];
        for (int t = 0; t < 27; ++t) c[t] = 0.0;
        const bool bnd = vi_bnd(g, i, j, k);
        c[13] = bnd ? g.db : g.d;
        if (!bnd) {
            if (!vi_bnd(g, i - 1, j, k)) c[12] = -g.sx;
            if (!vi_bnd(g, i + 1, j, k)) c[14] = -g.sx;
            if (g.dim >= 2) {
                if (!vi_bnd(g, i, j - 1, k)) c[10] = -g.sy;
                if (!vi_bnd(g, i, j + 1, k)) c[16] = -g.sy;
            }
            if (g.dim == 3) {
                if (!vi_bnd(g, i, j, k - 1)) c[4] = -g.sz;
                if (!vi_bnd(g, i, j, k + 1)) c[22] = -g.sz;
            }
        }
        for (int t = 0; t < 27; ++t) c[t] *= db;
        c[13] += da;
        const int z0 = g.dim == 3 ? 0 : 1, z1 = g.dim == 3 ? 2 : 1;
        const int y0 = g.dim >= 2 ? 0 : 1, y1 = g.dim >= 2 ? 2 : 1;
        for (int oz = z0; oz <= z1; ++oz)
            for (int oy = y0; oy <= y1; ++oy)
                for (int ox = 0; ox < 3; ++ox) {
                    const int t = 9 * oz + 3 * oy + ox;
                    g27[(long)t * g.n + q] = c[t];
                }
    }
}

// the reduced operator in 27-point form (coefficient c = 9(oz+1) + 3(oy+1)
// + (ox+1) of row q at g27[c n + q]): inactive rows keep the diagonal and
// the couplings to inactive interior neighbours; active rows are null
__global__ void __launch_bounds__(256)
k_vi_nullrow27(ViGrid g, const unsigned char* __restrict__ act, double* __restrict__ g27) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < g.n; q += (long)gridDim.x * 256L) {
        int i, j, k;
        vi_ijk(g, q, i, j, k);
        double c[27];
        for (int t = 0; t < 27; ++t) c[t] = 0.0;
        if (!act[q]) {
            const bool bnd = vi_bnd(g, i, j, k);
            c[13] = bnd ? g.db : g.d;
            if (!bnd) {
                const long sy = g.mx, sz = (long)g.mx * g.my;
                if (!vi_bnd(g, i - 1, j, k) && !act[q - 1]) c[12] = -g.sx;
                if (!vi_bnd(g, i + 1, j, k) && !act[q + 1]) c[14] = -g.sx;
                if (g.dim >= 2) {
                    if (!vi_bnd(g, i, j - 1, k) && !act[q - sy]) c[10] = -g.sy;
                    if (!vi_bnd(g, i, j + 1, k) && !act[q + sy]) c[16] = -g.sy;
                }
                if (g.dim == 3) {
                    if (!vi_bnd(g, i, j, k - 1) && !act[q - sz]) c[4] = -g.sz;
                    if (!vi_bnd(g, i, j, k + 1) && !act[q + sz]) c[22] = -g.sz;
                }
            }
        }
        // only the planes of in-grid offsets are ever read (9 in 2D, 3 in 1D)
        const int z0 = g.dim == 3 ? 0 : 1, z1 = g.dim == 3 ? 2 : 1;
        const int y0 = g.dim >= 2 ? 0 : 1, y1 = g.dim >= 2 ? 2 : 1;
        for (int oz = z0; oz <= z1; ++oz)
            for (int oy = y0; oy <= y1; ++oy)
                for (int ox = 0; ox < 3; ++ox) {
                    const int t = 9 * oz + 3 * oy + ox;
                    g27[(long)t * g.n + q] = c[t];
                }
    }
}

// x_q = 0 where the operator's diagonal is zero (an active vertex)
template <typename T>
__global__ void __launch_bounds__(256)
k_vi_mask(long n, const double* __restrict__ diag, T* __restrict__ x) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L)
        if (diag[q] == 0.0) x[q] = T(0);
}


// ------------------------------------------------------ minimal surface
//
// c/ch7/minimal.c:230-301 on the unit square (box stencil, 2D).  G holds
// the boundary values g at the boundary nodes (a node's own value, and the
// value minimal.c evaluates at a neighbour on the boundary: the same point
// for the power-of-two spacings of a refined DMDA).  No FMA contraction, so
// that the finite-difference Jacobian, which divides rounding by its step,
// follows the reference's arithmetic.

struct MsGrid {
    int mx, my;
    long n;
    double hx, hy, q;
};

// F at node (i, j) with u at node rp replaced by up (rp < 0: none)
__device__ __forceinline__ double ms_point(const MsGrid& g, const double* __restrict__ u,
                                           const double* __restrict__ G, int i, int j, long rp, double up) {
#pragma clang fp contract(off)
    const int mx = g.mx;
    const long q = (long)j * mx + i;
    auto U = [&](long p) { return p == rp ? up : u[p]; };
    if (j == 0 || i == 0 || i == mx - 1 || j == g.my - 1) return U(q) - G[q];
    const bool E = i + 1 == mx - 1, W = i - 1 == 0, N = j + 1 == g.my - 1, S = j - 1 == 0;
    const double uc = U(q);
    const double ue = E ? G[q + 1] : U(q + 1);
    const double uw = W ? G[q - 1] : U(q - 1);
    const double un = N ? G[q + mx] : U(q + mx);
    const double us = S ? G[q - mx] : U(q - mx);
    const double une = (E || N) ? G[q + mx + 1] : U(q + mx + 1);
    const double unw = (W || N) ? G[q + mx - 1] : U(q + mx - 1);
    const double use = (E || S) ? G[q - mx + 1] : U(q - mx + 1);
    const double usw = (W || S) ? G[q - mx - 1] : U(q - mx - 1);
    const double hx = g.hx, hy = g.hy;
    double dux, duy;
    dux = (ue - uc) / hx;
    duy = (un + une - us - use) / (4.0 * hy);
    const double De = pow(1.0 + (dux * dux + duy * duy), g.q);
    dux = (uc - uw) / hx;
    duy = (unw + un - usw - us) / (4.0 * hy);
    const double Dw = pow(1.0 + (dux * dux + duy * duy), g.q);
    dux = (ue + une - uw - unw) / (4.0 * hx);
    duy = (un - uc) / hy;
    const double Dn = pow(1.0 + (dux * dux + duy * duy), g.q);
    dux = (ue + use - uw - usw) / (4.0 * hx);
    duy = (uc - us) / hy;
    const double Ds = pow(1.0 + (dux * dux + duy * duy), g.q);
    return -(hy / hx) * (De * (ue - uc) - Dw * (uc - uw)) - (hx / hy) * (Dn * (un - uc) - Ds * (uc - us));
}

__global__ void __launch_bounds__(256)
k_ms_residual(MsGrid g, const double* __restrict__ u, const double* __restrict__ G, double* __restrict__ F) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < g.n; q += (long)gridDim.x * 256L)
        F[q] = ms_point(g, u, G, (int)(q % g.mx), (int)(q / g.mx), -1, 0.0);
}

// MatFDColoringApply with the "wp" step h = 1/d on the DMDA box coloring:
// each row's entry for a neighbour p is (F_q(u + h e_p) - F_q(u)) d, which
// is what the colored sweep computes (no other column of p's color lies in
// q's box).  A thread per row evaluates its nine perturbed residuals and
// writes the 2D stencil planes 9..17 of the 27-plane form.
__global__ void __launch_bounds__(256)
k_ms_fdjac(MsGrid g, const double* __restrict__ u, const double* __restrict__ G, const double* __restrict__ F0,
           double d, double h, double* __restrict__ g27) {
#pragma clang fp contract(off)
    const long n = g.n;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % g.mx), j = (int)(q / g.mx);
        const double f0 = F0[q];
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int c = 9 + (oy + 1) * 3 + (ox + 1);
                double v = 0.0;
                if (i + ox >= 0 && i + ox < g.mx && j + oy >= 0 && j + oy < g.my) {
                    const long p = q + (long)oy * g.mx + ox;
                    v = (ms_point(g, u, G, i, j, p, u[p] + h) - f0) * d;
                }
                g27[(long)c * n + q] = v;
            }
    }
}

// -snes_mf_operator: MatMult_MFFD's two vector steps around the residual
// evaluation, w = h a + U (VecWAXPY) and y = (F(w) - F(U)) (1/h) (VecAXPY,
// VecScale), without FMA contraction
__global__ void __launch_bounds__(256)
k_mffd_base(long n, double h, const double* __restrict__ a, const double* __restrict__ U, double* __restrict__ w) {
#pragma clang fp contract(off)
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) w[q] = h * a[q] + U[q];
}

__global__ void __launch_bounds__(256)
k_mffd_diff(long n, const double* __restrict__ F, double ih, double* __restrict__ y) {
#pragma clang fp contract(off)
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) y[q] = (y[q] - F[q]) * ih;
}

// A constant 2D 5-point operator (a Dirichlet row's diagonal db, an
// interior row's d and -sx, -sy toward interior neighbours: the rows
// Poisson2DJacobianLocal sets) in the 2D planes 9..17 of the 27-plane form
__global__ void __launch_bounds__(256)
k_const5_to_g27(int mx, int my, double d, double db, double sx, double sy, double* __restrict__ g27) {
    const long n = (long)mx * my;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % mx), j = (int)(q / mx);
        const bool in = i > 0 && i < mx - 1 && j > 0 && j < my - 1;
        for (int c = 9; c < 18; ++c) g27[(long)c * n + q] = 0.0;
        g27[13L * n + q] = in ? d : db;
        if (!in) continue;
        if (i - 1 > 0) g27[12L * n + q] = -sx;
        if (i + 1 < mx - 1) g27[14L * n + q] = -sx;
        if (j - 1 > 0) g27[10L * n + q] = -sy;
        if (j + 1 < my - 1) g27[16L * n + q] = -sy;
    }
}

// The output policy of a stencil pass applied to a product A u formed
// elsewhere (a matrix-free operator): Out(u, Au) at every point, u or Au
// null meaning zero; the Jacobi factor from the diagonal of the 27-plane
// preconditioning operator g.
template <typename T, class Out>
__global__ void __launch_bounds__(256)
k_out_au(long n, const T* __restrict__ u, const T* __restrict__ Au, const double* __restrict__ g, int jac, Out out,
         double* __restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const T uc = u ? u[q] : T(0);
        const T au = Au ? Au[q] : T(0);
        const T dq = T(g[13 * n + q]);
        const T idg = jac ? (dq != T(0) ? T(1) / dq : T(0)) : T(1);
        out.point(q, uc, au, idg, a0, a1);
    }
    block_partials<Out::NDOT, 4>(a0, a1, part, blockIdx.x);
}

// coarse (I, J) <- fine (2I, 2J)
__global__ void __launch_bounds__(256)
k_inject2(int mxc, int myc, int mxf, const double* __restrict__ xf, double* __restrict__ xc) {
    const long nc = (long)mxc * myc;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < nc; q += (long)gridDim.x * 256L) {
        const int I = (int)(q % mxc), J = (int)(q / mxc);
        xc[q] = xf[(long)(2 * J) * mxf + 2 * I];
    }
}

// per block: max |a(q, p) - a(p, q)| over the 2D box entries (MatIsSymmetric)
__global__ void __launch_bounds__(256)
k_sym27(int mx, int my, const double* __restrict__ g27, double* __restrict__ part) {
    const long n = (long)mx * my;
    double m = 0.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        const int i = (int)(q % mx), j = (int)(q / mx);
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                if (i + ox < 0 || i + ox >= mx || j + oy < 0 || j + oy >= my) continue;
                const long p = q + (long)oy * mx + ox;
                const double a = g27[(long)(9 + (oy + 1) * 3 + (ox + 1)) * n + q];
                const double b = g27[(long)(9 + (1 - oy) * 3 + (1 - ox)) * n + p];
                m = fmax(m, fabs(a - b));
            }
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// ------------------------------------------------------------ GMRES
//
// V holds the Krylov basis as m contiguous vectors of length n.

// per block: (V_j, w) for j < m, written to part[j * gridDim.x + block]
__global__ void __launch_bounds__(256)
k_mdot(long n, int m, const double* __restrict__ V, const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double sm[4];
    for (int j = 0; j < m; ++j) {
        const double* v = V + (long)j * n;
        double a = 0.0;
        for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) a += v[q] * w[q];
        a = wave_sum(a);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) part[(long)j * gridDim.x + blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
        __syncthreads();
    }
}

// out[j] = sum_b part[j * nb + b], one block per j, a fixed order
__global__ void __launch_bounds__(256)
k_mfinal(int nb, const double* __restrict__ part, double* __restrict__ out) {
    const double* p = part + (long)blockIdx.x * nb;
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) a += p[b];
    __shared__ double sm[4];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

// w += s * sum_{j < m} c_j V_j   (c on the device)
__global__ void __launch_bounds__(256)
k_maxpy(long n, int m, double s, const double* __restrict__ c, const double* __restrict__ V, double* __restrict__ w) {
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
        double a = 0.0;
        for (int j = 0; j < m; ++j) a += c[j] * V[(long)j * n + q];
        w[q] += s * a;
    }
}

// v = w / sqrt(*nn) (nn: a device scalar holding |w|^2), or a copy when 0
__global__ void __launch_bounds__(256)
k_normalize(long n, const double* __restrict__ nn, const double* __restrict__ w, double* __restrict__ v) {
    const double t = sqrt(*nn);
    const double s = t != 0.0 ? 1.0 / t : 1.0;
    for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) v[q] = w[q] * s;
}
}  // namespace p4g
