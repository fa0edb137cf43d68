// p4g_batch.hpp -- ragged batched launches for the serving loop.
//
// One launch covers a list of work items (a grid function of one request on
// one multigrid level); items may have different sizes.  Item t owns the
// blocks [start_t, start_t + gx*gy*gz) of a 1-D launch grid; a block finds
// its item by binary search over the starts and decodes its (bx,by,bz)
// within the item's own grid, then runs exactly the per-block code of the
// single-grid kernels (p4g_kernels.hpp), so a request's arithmetic does not
// depend on which other requests share the launch.
#pragma once

#include "p4g_kernels.hpp"

namespace p4g {

struct GridRef {
    int  mx, my, mz, kch;   // grid and marching chunk
    int  gx, gy, gz;        // the item's block grid
    long start;             // first block of the item in the launch
};

template <class Item>
__device__ __forceinline__ int find_item(const Item* __restrict__ it, int n, long b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (it[mid].g.start <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void decode(const GridRef& g, long b, int& bx, int& by, int& bz, long& lb) {
    lb = b - g.start;
    long r = lb;
    bx = (int)(r % g.gx);
    r /= g.gx;
    by = (int)(r % g.gy);
    bz = (int)(r / g.gy);
}

// ---------------------------------------------------------------- stencil

template <class Op, class Src, class Out>
struct StencilItem {
    GridRef g;
    double* part;          // this item's partial sums (2 per block)
    const double* flag;    // optional: skip the item when *flag == 0
    Op op;
    Src src;
    Out out;
};

template <typename T, int DIM, class Op, class Src, class Out>
__global__ void __launch_bounds__(NT)
k_stencil_batch(const StencilItem<Op, Src, Out>* __restrict__ items, int nitems) {
    const int t = find_item(items, nitems, (long)blockIdx.x);
    const StencilItem<Op, Src, Out>& it = items[t];
    if (it.flag && *it.flag == 0.0) return;   // whole block, uniform
    int bx, by, bz;
    long lb;
    decode(it.g, blockIdx.x, bx, by, bz, lb);
    double a0 = 0.0, a1 = 0.0;
    stencil_body<T, DIM, Op, Src, Out>(bx, by, bz, it.g.mx, it.g.my, it.g.mz, it.g.kch, it.op, it.src, it.out,
                                       a0, a1);
    block_partials<Out::NDOT>(a0, a1, it.part, lb);
}

// y = c1 u + c2 D^-1 (b - A u); with dots (runtime flag): acc0 += y.b, acc1 += y.y
template <typename T>
struct OutChebR {
    const T* __restrict__ b;
    T* __restrict__ y;
    T c1, c2;
    int dots;
    static constexpr int NDOT = 2, NPRE = 1;
    template <int V>
    __device__ __forceinline__ void pre(long q, TV<T, V>* p) const { p[0] = ldv<T, V>(b + q); }
    __device__ __forceinline__ void pre1(long q, T* p) const { p[0] = b[q]; }
    __device__ __forceinline__ void pointp(long q, T u, T Au, T idg, const T* p, double& a0, double& a1) const {
        const T v = c1 * u + c2 * (idg * (p[0] - Au));
        y[q] = v;
        if (dots) {
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
        stv<T, V>(y + q, v);
        if (dots)
#pragma unroll
            for (int e = 0; e < V; ++e) {
                a0 += double(v.v[e]) * double(p[0].v[e]);
                a1 += double(v.v[e]) * double(v.v[e]);
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

// ---------------------------------------------------------------- transfers

template <typename T>
struct RestrictItem {
    GridRef g;             // coarse grid
    int mxf, myf, mzf;
    const T* rf;
    T* bc;
};

template <typename T, int DIM>
__global__ void __launch_bounds__(NT)
k_restrict_batch(const RestrictItem<T>* __restrict__ items, int nitems) {
    const int t = find_item(items, nitems, (long)blockIdx.x);
    const RestrictItem<T>& it = items[t];
    int bx, by, bz;
    long lb;
    decode(it.g, blockIdx.x, bx, by, bz, lb);
    restrict_body<T, DIM>(bx, by, bz, it.mxf, it.myf, it.mzf, it.g.mx, it.g.my, it.g.mz, it.g.kch, it.rf, it.bc);
}

// prolongation fused with the first smoothing step (zero-guess base c b)
template <typename T, class Op, class Out>
struct ProlongSmoothItem {
    GridRef g;             // fine grid, stencil launch geometry
    int mxc, myc;          // coarse grid
    const T* b;            // base: t = c b + P x_c
    T c;
    const T* xc;
    double* part;
    Op op;
    Out out;
};

template <typename T, int DIM, class Op, class Out>
__global__ void __launch_bounds__(NT)
k_prolong_smooth_batch(const ProlongSmoothItem<T, Op, Out>* __restrict__ items, int nitems) {
    const int t = find_item(items, nitems, (long)blockIdx.x);
    const ProlongSmoothItem<T, Op, Out>& it = items[t];
    int bx, by, bz;
    long lb;
    decode(it.g, blockIdx.x, bx, by, bz, lb);
    double a0 = 0.0, a1 = 0.0;
    prolong_smooth_body<T, DIM, Op, true, Out>(bx, by, bz, it.g.mx, it.g.my, it.g.mz, it.mxc, it.myc, it.g.kch,
                                               it.op, it.b, it.c, it.xc, it.out, a0, a1);
    block_partials<Out::NDOT>(a0, a1, it.part, lb);
}

// ------------------------------------------------------------- coarse solve

template <typename T>
struct CoarseItem {
    int n;
    const double* Ainv;    // dense inverse of the coarsest operator
    const T* b;
    T* x;
};

// one block of FT threads per item
template <typename T>
__global__ void __launch_bounds__(FT)
k_coarse_batch(const CoarseItem<T>* __restrict__ items) {
    const CoarseItem<T>& it = items[blockIdx.x];
    coarse_rows<T>(it.n, it.Ainv, it.b, it.x, threadIdx.x >> 6, FT / 64);
}

// ---------------------------------------------------------- vector kernels

// x += a p in FP64 ; r -= a w in FP32 ; acc0 += r.r  (a = S[S_ALPHA])
struct UpdateItem {
    GridRef g;             // only start and gx (= number of blocks) are used
    long n;
    const double* S;
    double* x;
    float* r;
    const float *p, *w;
    double* part;
};

__global__ void __launch_bounds__(256)
k_update_batch(const UpdateItem* __restrict__ items, int nitems) {
    const int t = find_item(items, nitems, (long)blockIdx.x);
    const UpdateItem& it = items[t];
    const long lb = blockIdx.x - it.g.start;
    const double a = it.S[S_ALPHA];
    const float af = float(a);
    double acc = 0.0;
    for (long q = lb * 256 + threadIdx.x; q < it.n; q += (long)it.g.gx * 256) {
        it.x[q] += a * double(it.p[q]);
        const float rv = it.r[q] - af * it.w[q];
        it.r[q] = rv;
        acc += double(rv) * double(rv);
    }
    __shared__ double sm[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        it.part[2 * lb] = sm[0] + sm[1] + sm[2] + sm[3];
        it.part[2 * lb + 1] = 0.0;
    }
}

// -------------------------------------------------------------- finalisation

struct FinItem {
    const double* part;
    long nparts;
    double* S;
    int mode, first;
    const double* flag;    // optional: skip when *flag == 0
};

// launch: <<<nitems, FT>>>
__global__ void __launch_bounds__(FT)
k_finalize_batch(const FinItem* __restrict__ items) {
    const FinItem& it = items[blockIdx.x];
    if (it.flag && *it.flag == 0.0) return;
    double A, B;
    sum_parts(it.part, it.nparts, A, B);
    if (threadIdx.x == 0) finalize_scalars(it.S, it.mode, it.first, A, B);
}

}  // namespace p4g
