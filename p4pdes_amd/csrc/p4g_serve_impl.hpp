// p4g_serve_impl.hpp -- continuous-batching solve server (included by
// p4g_solver.hip, which defines the solver types it builds on).
//
// A worker thread owns one HIP stream.  Each round advances every active
// request by one PCG iteration of the mixed-precision solver (FP32 vectors
// and V-cycle, FP64 iterate and residual replacement; the same algorithm as
// -p4g_precision mixed): one ragged-batched launch per phase covers all
// requests, and the V-cycle batches equal-sized levels of all requests
// (descending by grid size, coarsest solves, ascending).  The V-cycle launch
// sequence is captured in a hipGraph and replayed every round until the
// active set changes.  Requests are admitted between rounds, in arrival
// order, when their device footprint fits the HBM budget; finished requests
// release theirs.  Each request's arithmetic is independent of the others
// (per-request blocks and partial sums), so its result is the same alone
// and in a mix.  Nothing slow runs in the worker's loop: a request's device
// vectors come from a per-solver pool (no hipMalloc/hipFree once warm), its
// right-hand side is copied into pooled pinned host memory by the submitting
// thread, and its result goes back through that pinned buffer, copied into
// the caller's array by a completion thread once the device copy's event has
// fired.

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "p4g_batch.hpp"

namespace p4g {

struct SrvLevel {
    float *b = nullptr, *x = nullptr, *r = nullptr;
};

enum { RQ_QUEUED = 0, RQ_INIT = 1, RQ_ITER = 2, RQ_DONE = 3, RQ_FAILED = -1 };

struct SrvReq {
    long id = 0;
    p4g_solver_s* sol = nullptr;
    const double* hb = nullptr;
    double* hx = nullptr;
    int dim = 3, nlev = 0;
    long n = 0;
    size_t bytes = 0;
    int slot = -1;
    double *b64 = nullptr, *x64 = nullptr;
    float *r32 = nullptr, *z32 = nullptr, *pa = nullptr, *pb = nullptr, *w32 = nullptr;
    std::vector<SrvLevel> lv;
    double* part = nullptr;
    long npart = 0;
    std::vector<void*> allocs;
    int state = RQ_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int its = 0, nrep = 0, reason = 0;
    bool replace_now = false, replaced = false;
    double bnorm = 0, target = 0, rr_last = 0, dpi_old = 0, rn = 0, rz_old = 0;
    float *pold = nullptr, *pnew = nullptr;
    std::string err;
    double* S = nullptr;   // device scalar block (32 doubles) of the slot
    std::vector<size_t> layout;    // byte sizes of allocs, in order (the pool's key)
    double* pin = nullptr;         // pinned host staging: b at submit, x at the end
    hipEvent_t ev = nullptr;       // the result's device-to-host copy
    hipEvent_t ev_in = nullptr;    // b's upload
    hipEvent_t ev_k = nullptr;     // the request's last kernel
};

// a launch of a batched kernel over a list of items
struct BLaunch {
    int kind = 0;          // which kernel
    int dim = 3;
    size_t off = 0;        // byte offset of the items in the device buffer
    int nitems = 0;
    long blocks = 0;
    size_t len = 0;        // bytes of the items
};

}  // namespace p4g

struct p4g_server_s {
    size_t budget = 0, used = 0, peak = 0;
    bool use_graph = true;
    int max_slots = 64;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<p4g::SrvReq*> queue;
    std::vector<p4g::SrvReq*> active;
    std::map<long, p4g::SrvReq*> all;
    std::vector<int> free_slots;
    long next_id = 1;
    bool stop = false;
    int device = 0;
    std::thread th;
    hipStream_t st = nullptr;
    double *S = nullptr, *hS = nullptr;
    char *d_round = nullptr, *h_round = nullptr;
    size_t round_cap = 0, round_bytes = 0;
    char *d_vc = nullptr, *h_vc = nullptr;
    size_t vc_cap = 0, vc_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<long> graph_key;          // the active set the graph was captured for
    std::vector<long> vc_key;             // the active set the V-cycle tables were built for
    std::vector<p4g::BLaunch> vseq;       // their launch sequence
    bool vc_built = false;
    int vc_stable = 0;                    // rounds the active set has held
    int graph_after = 32;                 // -p4g_serve_graph_after: rounds before a capture
    // pools: idle device vector sets per layout (the byte sizes of a
    // request's vectors in allocation order), pinned staging buffers per
    // length; pool_bytes counts the idle device sets against the budget
    // together with used
    std::map<std::vector<size_t>, std::vector<std::vector<void*>>> dpool;
    std::map<long, std::vector<double*>> ppool;
    // a pinned staging arena allocated with the server (-p4g_serve_staging_mb,
    // default 1024): first-time staging buffers are carved from it instead of
    // pinned one by one (pinning 100+ MB takes tens of ms); free ranges by offset
    char* arena = nullptr;
    size_t arena_size = 0;
    std::map<size_t, size_t> arena_free;
    size_t pool_bytes = 0;
    // requests whose result is being copied back; the completion thread
    std::deque<p4g::SrvReq*> copying;
    std::condition_variable cv_copy;
    bool stop_copy = false;
    std::vector<std::thread> th_copy;     // completion threads
    hipStream_t cst = nullptr;            // uploads and result copies
    std::vector<p4g::SrvReq*> loading;    // admitted, b still uploading (worker thread only)
    // statistics
    std::vector<double> lat_ms, queue_ms;
    long submitted = 0, completed = 0, rejected = 0, rounds = 0, captures = 0;
    double t_first = -1, t_last = 0, dof_done = 0;
    std::string fatal;
};

namespace p4g {

static double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// footprint of one request on the device (bytes)
static size_t req_bytes(p4g_solver_s* s) {
    const LevelInfo& F = s->lv.back();
    const long n = F.n;
    size_t b = 2 * n * sizeof(double) + 6 * n * sizeof(float);   // 5 Krylov vectors + fine-level r
    for (int l = 0; l + 1 < s->nlev; ++l) b += (l > 0 ? 3 : 2) * s->lv[l].n * sizeof(float);
    const long nb = std::max<long>(max_part_blocks(s->lv), vblocks(n));
    b += 2 * nb * sizeof(double);
    return b;
}

static int serve_supported(p4g_solver_s* s) {
    if (s->pctype != PCT_MG || s->nlev < 2)
        return set_error(PETSC_ERR_SUP, "serving needs -pc_type mg with at least 2 levels");
    if (s->nu != 1 || s->cycles != 1)
        return set_error(PETSC_ERR_SUP, "serving runs V(1,1) cycles: -mg_levels_ksp_max_it 1, -pc_mg_cycle_type v");
    if (s->lv.back().dim < 2)
        return set_error(PETSC_ERR_SUP, "serving supports 2D and 3D grids");
    for (const LevelInfo& L : s->lv)
        if (L.var || L.d != L.db)
            return set_error(PETSC_ERR_SUP, "serving supports constant-coefficient stencils");
    if (s->smoother != SM_CHEB && s->smoother != SM_RICH)
        return set_error(PETSC_ERR_SUP, "serving needs a Chebyshev or Richardson smoother");
    if (s->galerkin || s->fmg || s->kaskade || s->mcycles != 1)
        return set_error(PETSC_ERR_SUP, "serving runs rediscretized multiplicative V-cycles (no -pc_mg_galerkin, "
                                        "-pc_mg_type full|kaskade or -pc_mg_multiplicative_cycles)");
    if (s->sor) return set_error(PETSC_ERR_SUP, "serving runs Jacobi-preconditioned smoothers (no SOR)");
    if (s->coarse_its > 0)
        return set_error(PETSC_ERR_SUP, "serving solves the coarsest level directly (no -mg_coarse_ksp_type cg)");
    if (s->ksp_rich || s->ksp_gmres || s->vi_mode)
        return set_error(PETSC_ERR_SUP, "serving runs CG solves (no -ksp_type richardson|gmres or VI solver)");
    return 0;
}

static double level_scale(p4g_solver_s* s, const LevelInfo& L) {
    return s->smoother == SM_CHEB ? 2.0 / (L.emax + L.emin) : s->rich_scale;
}

static OpConst<float> op32(p4g_solver_s* s, const LevelInfo& L) {
    OpConst<float> o;
    o.d = float(L.d); o.db = float(L.db); o.sx = float(L.sx); o.sy = float(L.sy); o.sz = float(L.sz);
    o.id = s->jacobi ? float(1.0 / L.d) : 1.0f;
    o.idb = s->jacobi ? float(1.0 / L.db) : 1.0f;
    o.diff = s->fp32_diff && std::fabs(L.d - 2.0 * (L.sx + L.sy + L.sz)) <= 1e-12 * std::fabs(L.d) ? 1 : 0;
    return o;
}

static OpConst<double> op64(const LevelInfo& L) {
    OpConst<double> o;
    o.d = L.d; o.db = L.db; o.sx = L.sx; o.sy = L.sy; o.sz = L.sz; o.id = o.idb = 1.0;
    return o;
}

static const double* solver_ainv(p4g_solver_s* s) {
    return s->e32 ? s->e32->Ainv : (s->e64 ? s->e64->Ainv : nullptr);
}

// a stencil launch over level L with V x points per lane (V = 2: the pair
// geometry of the fused prolongation + smoothing kernel)
static GridRef gref_stencil(const LevelInfo& L, long start, int V) {
    const dim3 g = stencil_grid_v(L.dim, L.mx, L.my, L.mz, V);
    GridRef r;
    r.mx = L.mx; r.my = L.my; r.mz = L.mz; r.kch = kchf(L.dim, L.my, L.mz);
    r.gx = g.x; r.gy = g.y; r.gz = g.z; r.start = start;
    return r;
}

// the fused prolongation + smoothing launch over level L
static GridRef gref_fused(const LevelInfo& L, long start) {
    const dim3 g = fused_grid(L.dim, L.mx, L.my, L.mz);
    GridRef r;
    r.mx = L.mx; r.my = L.my; r.mz = L.mz; r.kch = kch_fused(L.dim, L.my, L.mz);
    r.gx = g.x; r.gy = g.y; r.gz = g.z; r.start = start;
    return r;
}

// restriction launch over the coarse grid L (lanes over its columns 1..mx-1)
static GridRef gref_restrict(const LevelInfo& L, long start) {
    const int kch = kch_restrict(L.dim, L.my, L.mz);
    const dim3 g = transfer_grid(L.dim, L.mx, L.my, L.mz, kch);
    GridRef r;
    r.mx = L.mx; r.my = L.my; r.mz = L.mz; r.kch = kch;
    r.gx = g.x; r.gy = g.y; r.gz = g.z; r.start = start;
    return r;
}

static long gblocks(const GridRef& g) { return (long)g.gx * g.gy * g.gz; }

// ---------------------------------------------------------- item packing

struct Packer {
    std::vector<char> buf;
    size_t put(const void* p, size_t bytes) {
        size_t off = (buf.size() + 255) & ~(size_t)255;
        buf.resize(off + bytes);
        std::memcpy(buf.data() + off, p, bytes);
        return off;
    }
};


enum { BK_DIR = 0, BK_FIN, BK_UPD, BK_REP, BK_RESID, BK_RESTRICT, BK_PSMOOTH, BK_COARSE };

typedef StencilItem<OpConst<float>, SrcDir<float>, OutApplyDot<float>> DirItem;
typedef StencilItem<OpConst<double>, SrcArr<double>, OutResidCvt<double, float>> RepItem;
typedef StencilItem<OpConst<float>, SrcArr<float>, OutScaleU<float, OutResid<float, 0>>> ResidItem;
typedef ProlongSmoothItem<float, OpConst<float>, OutChebR<float>> PSmoothItem;

static int blaunch(p4g_server_s* v, const BLaunch& L, const char* base, size_t table_bytes) {
    if (L.nitems == 0 || L.blocks == 0) return 0;
    if (!base || L.off + L.len > table_bytes)
        return set_error(PETSC_ERR_LIB, "internal: batched launch outside its item table (%zu + %zu > %zu)", L.off,
                         L.len, table_bytes);
    const void* items = base + L.off;
    hipStream_t st = v->st;
    const dim3 g((unsigned)L.blocks);
    switch (L.kind) {
    case BK_DIR:
        if (L.dim == 3) k_stencil_batch<float, 3, OpConst<float>, SrcDir<float>, OutApplyDot<float>><<<g, dim3(TX, TY), 0, st>>>((const DirItem*)items, L.nitems);
        else k_stencil_batch<float, 2, OpConst<float>, SrcDir<float>, OutApplyDot<float>><<<g, dim3(TX, TY), 0, st>>>((const DirItem*)items, L.nitems);
        break;
    case BK_REP:
        if (L.dim == 3) k_stencil_batch<double, 3, OpConst<double>, SrcArr<double>, OutResidCvt<double, float>><<<g, dim3(TX, TY), 0, st>>>((const RepItem*)items, L.nitems);
        else k_stencil_batch<double, 2, OpConst<double>, SrcArr<double>, OutResidCvt<double, float>><<<g, dim3(TX, TY), 0, st>>>((const RepItem*)items, L.nitems);
        break;
    case BK_RESID:
        if (L.dim == 3) k_stencil_batch<float, 3, OpConst<float>, SrcArr<float>, OutScaleU<float, OutResid<float, 0>>><<<g, dim3(TX, TY), 0, st>>>((const ResidItem*)items, L.nitems);
        else k_stencil_batch<float, 2, OpConst<float>, SrcArr<float>, OutScaleU<float, OutResid<float, 0>>><<<g, dim3(TX, TY), 0, st>>>((const ResidItem*)items, L.nitems);
        break;
    case BK_PSMOOTH:
        if (L.dim == 3) k_prolong_smooth_batch<float, 3, OpConst<float>, OutChebR<float>><<<g, dim3(TX, TY), 0, st>>>((const PSmoothItem*)items, L.nitems);
        else k_prolong_smooth_batch<float, 2, OpConst<float>, OutChebR<float>><<<g, dim3(TX, TY), 0, st>>>((const PSmoothItem*)items, L.nitems);
        break;
    case BK_RESTRICT:
        if (L.dim == 3) k_restrict_batch<float, 3><<<g, dim3(TX, TY), 0, st>>>((const RestrictItem<float>*)items, L.nitems);
        else k_restrict_batch<float, 2><<<g, dim3(TX, TY), 0, st>>>((const RestrictItem<float>*)items, L.nitems);
        break;
    case BK_UPD:
        k_update_batch<<<g, 256, 0, st>>>((const UpdateItem*)items, L.nitems);
        break;
    case BK_FIN:
        k_finalize_batch<<<dim3((unsigned)L.nitems), FT, 0, st>>>((const FinItem*)items);
        break;
    case BK_COARSE:
        k_coarse_batch<float><<<dim3((unsigned)L.nitems), FT, 0, st>>>((const CoarseItem<float>*)items);
        break;
    }
    LAUNCH_CHECK();
    return 0;
}

static int upload(p4g_server_s* v, const Packer& pk, char** dbuf, char** hbuf, size_t* cap, size_t* used) {
    const size_t bytes = pk.buf.size();
    *used = bytes;
    if (bytes == 0) return 0;
    // work queued earlier may still read the device table or the staging buffer
    HIP_TRY(hipStreamSynchronize(v->st));
    if (bytes > *cap) {
        if (*dbuf) hipFree(*dbuf);
        if (*hbuf) hipHostFree(*hbuf);
        *cap = std::max<size_t>(bytes * 2, 1 << 16);
        HIP_TRY(hipMalloc((void**)dbuf, *cap));
        HIP_TRY(hipHostMalloc((void**)hbuf, *cap, hipHostMallocDefault));
    }
    std::memcpy(*hbuf, pk.buf.data(), bytes);
    HIP_TRY(hipMemcpyAsync(*dbuf, *hbuf, bytes, hipMemcpyHostToDevice, v->st));
    return 0;
}

// ------------------------------------------------------------- V-cycle plan

// Batched V(1,1) over all active requests: equal-sized levels share a
// launch.  Descending sizes: zero-guess smoothing + residual, restriction;
// all coarsest solves; ascending sizes: prolongation (+ c b), smoothing.
static int build_vcycle(p4g_server_s* v, Packer& pk, std::vector<BLaunch>& seq) {
    if (!seq.empty()) return set_error(PETSC_ERR_LIB, "internal: V-cycle plan built onto a non-empty sequence");
    struct LvRef { SrvReq* r; int l; };
    std::map<std::pair<int, int>, std::vector<LvRef>, std::greater<std::pair<int, int>>> groups;
    for (SrvReq* r : v->active)
        for (int l = 1; l < r->nlev; ++l) groups[{r->sol->lv[l].mx, r->dim}].push_back({r, l});
    auto bvec = [](SrvReq* r, int l) -> float* { return l == r->nlev - 1 ? r->r32 : r->lv[l].b; };
    // descent
    for (auto& kv : groups) {
        const int dim = kv.first.second;
        std::vector<ResidItem> ri;
        std::vector<RestrictItem<float>> rs;
        long s1 = 0, s2 = 0;
        for (const LvRef& q : kv.second) {
            p4g_solver_s* s = q.r->sol;
            const LevelInfo& L = s->lv[q.l];
            const LevelInfo& C = s->lv[q.l - 1];
            ResidItem it;
            it.g = gref_stencil(L, s1, VecOf<float>::V);
            s1 += gblocks(it.g);
            it.part = nullptr;
            it.flag = nullptr;
            it.op = op32(s, L);
            const float c = float(level_scale(s, L) * (s->jacobi ? 1.0 / L.d : 1.0));
            it.src = SrcArr<float>{bvec(q.r, q.l)};
            it.out = OutScaleU<float, OutResid<float, 0>>{c, OutResid<float, 0>{bvec(q.r, q.l), q.r->lv[q.l].r}};
            ri.push_back(it);
            RestrictItem<float> rt;
            rt.g = gref_restrict(C, s2);
            s2 += gblocks(rt.g);
            rt.mxf = L.mx; rt.myf = L.my; rt.mzf = L.mz;
            rt.rf = q.r->lv[q.l].r;
            rt.bc = q.r->lv[q.l - 1].b;
            rs.push_back(rt);
        }
        BLaunch a; a.kind = BK_RESID; a.dim = dim; a.nitems = (int)ri.size(); a.blocks = s1;
        a.off = pk.put(ri.data(), ri.size() * sizeof(ResidItem)); a.len = ri.size() * sizeof(ResidItem);
        seq.push_back(a);
        BLaunch b; b.kind = BK_RESTRICT; b.dim = dim; b.nitems = (int)rs.size(); b.blocks = s2;
        b.off = pk.put(rs.data(), rs.size() * sizeof(RestrictItem<float>)); b.len = rs.size() * sizeof(RestrictItem<float>);
        seq.push_back(b);
    }
    // coarsest solves
    {
        std::vector<CoarseItem<float>> ci;
        for (SrvReq* r : v->active) {
            CoarseItem<float> it;
            it.n = (int)r->sol->lv[0].n;
            it.Ainv = solver_ainv(r->sol);
            it.b = r->lv[0].b;
            it.x = r->lv[0].x;
            ci.push_back(it);
        }
        BLaunch c; c.kind = BK_COARSE; c.nitems = (int)ci.size(); c.blocks = c.nitems;
        c.off = pk.put(ci.data(), ci.size() * sizeof(CoarseItem<float>)); c.len = ci.size() * sizeof(CoarseItem<float>);
        seq.push_back(c);
    }
    // ascent: prolongation fused with the post-smoothing step
    for (auto it = groups.rbegin(); it != groups.rend(); ++it) {
        const int dim = it->first.second;
        std::vector<PSmoothItem> pi;
        long s1 = 0;
        for (const LvRef& q : it->second) {
            p4g_solver_s* s = q.r->sol;
            const LevelInfo& L = s->lv[q.l];
            const LevelInfo& C = s->lv[q.l - 1];
            const double sc = level_scale(s, L);
            const bool fin = q.l == q.r->nlev - 1;
            PSmoothItem m;
            m.g = gref_fused(L, s1);
            s1 += gblocks(m.g);
            m.mxc = C.mx;
            m.myc = C.my;
            m.b = bvec(q.r, q.l);
            m.c = float(sc * (s->jacobi ? 1.0 / L.d : 1.0));
            m.xc = q.r->lv[q.l - 1].x;
            m.part = q.r->part;
            m.op = op32(s, L);
            m.out = OutChebR<float>{bvec(q.r, q.l), fin ? q.r->z32 : q.r->lv[q.l].x, 1.0f, float(sc), fin ? 1 : 0};
            pi.push_back(m);
        }
        BLaunch a; a.kind = BK_PSMOOTH; a.dim = dim; a.nitems = (int)pi.size(); a.blocks = s1;
        a.off = pk.put(pi.data(), pi.size() * sizeof(PSmoothItem)); a.len = pi.size() * sizeof(PSmoothItem);
        seq.push_back(a);
    }
    return 0;
}

// ---------------------------------------------------------- request memory

static void req_free(SrvReq* r) {
    for (void* p : r->allocs) hipFree(p);
    r->allocs.clear();
}

// a pinned staging buffer of n doubles (pooled per length, else carved
// from the arena, else pinned on its own); call without the lock
static int pin_get(p4g_server_s* v, long n, double** out) {
    {
        std::lock_guard<std::mutex> lk(v->mu);
        std::vector<double*>& q = v->ppool[n];
        if (!q.empty()) {
            *out = q.back();
            q.pop_back();
            return 0;
        }
        const size_t want = ((std::max<long>(n, 1) * sizeof(double)) + 4095) & ~(size_t)4095;
        for (auto it = v->arena_free.begin(); it != v->arena_free.end(); ++it) {
            if (it->second < want) continue;
            const size_t off = it->first, len = it->second;
            v->arena_free.erase(it);
            if (len > want) v->arena_free[off + want] = len - want;
            *out = (double*)(v->arena + off);
            return 0;   // the buffer stays with its length class (ppool) once released
        }
    }
    if (hipHostMalloc((void**)out, std::max<long>(n, 1) * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return set_error(PETSC_ERR_MEM, "serving: pinned staging of %ld doubles", n);
    return 0;
}

// the byte sizes of a request's device vectors in allocation order: b64,
// x64, the five FP32 Krylov vectors, each level's b, x (below the finest)
// and r (above the coarsest), the partial sums
static std::vector<size_t> req_layout(p4g_solver_s* s) {
    const long n = s->lv.back().n;
    std::vector<size_t> z = {size_t(n) * 8, size_t(n) * 8};
    for (int q = 0; q < 5; ++q) z.push_back(size_t(n) * 4);
    for (int l = 0; l < s->nlev; ++l) {
        const size_t nl = size_t(s->lv[l].n);
        if (l + 1 < s->nlev) { z.push_back(nl * 4); z.push_back(nl * 4); }
        if (l > 0) z.push_back(nl * 4);
    }
    z.push_back(2 * size_t(std::max<long>(max_part_blocks(s->lv), vblocks(n))) * 8);
    return z;
}

static int req_admit(p4g_server_s* v, SrvReq* r) {
    p4g_solver_s* s = r->sol;
    const LevelInfo& F = s->lv.back();
    const long n = F.n;
    r->nlev = s->nlev;
    r->dim = F.dim;
    r->n = n;
    r->layout = req_layout(s);
    // an idle vector set of this exact layout, else new allocations (idle
    // sets of other layouts released first when the budget needs their room)
    std::vector<void*> set;
    std::vector<void*> evict;
    {
        std::lock_guard<std::mutex> lk(v->mu);
        auto it = v->dpool.find(r->layout);
        if (it != v->dpool.end() && !it->second.empty()) {
            set = std::move(it->second.back());
            it->second.pop_back();
            v->pool_bytes -= r->bytes;
        } else {
            for (auto& kv : v->dpool) {
                size_t each = 0;
                for (size_t b : kv.first) each += b;
                while (!kv.second.empty() && v->used + v->pool_bytes > v->budget) {
                    for (void* p : kv.second.back()) evict.push_back(p);
                    kv.second.pop_back();
                    v->pool_bytes -= each;
                }
            }
        }
    }
    for (void* p : evict) hipFree(p);
    if (set.empty()) {
        for (size_t b : r->layout) {
            void* p = nullptr;
            const hipError_t e = hipMalloc(&p, std::max<size_t>(b, 8));
            if (e != hipSuccess) {
                for (void* q : set) hipFree(q);
                return set_error(PETSC_ERR_MEM, "serving: hipMalloc(%zu): %s", b, hipGetErrorString(e));
            }
            set.push_back(p);
        }
    }
    if (set.size() != r->layout.size()) return set_error(PETSC_ERR_LIB, "internal: request vector set of another layout");
    r->allocs = set;
    size_t k = 0;
    r->b64 = (double*)set[k++];
    r->x64 = (double*)set[k++];
    float** f32[5] = {&r->r32, &r->z32, &r->pa, &r->pb, &r->w32};
    for (float** q : f32) *q = (float*)set[k++];
    r->lv.assign(r->nlev, SrvLevel());
    for (int l = 0; l < r->nlev; ++l) {
        if (l + 1 < r->nlev) {
            r->lv[l].b = (float*)set[k++];
            r->lv[l].x = (float*)set[k++];
        }
        if (l > 0) r->lv[l].r = (float*)set[k++];
    }
    r->npart = std::max<long>(max_part_blocks(s->lv), vblocks(n));
    r->part = (double*)set[k++];
    r->slot = v->free_slots.back();
    v->free_slots.pop_back();
    r->S = v->S + 32 * r->slot;
    // b (staged by submit) on the copy stream; the request joins the rounds
    // once this upload's event has fired (req_activate)
    HIP_TRY(hipMemcpyAsync(r->b64, r->pin, n * 8, hipMemcpyHostToDevice, v->cst));
    HIP_TRY(hipEventRecord(r->ev_in, v->cst));
    return 0;
}

// the initial state once b is on the device: x = 0, p = 0, r32 = b, |b|^2
// -> outer TMP0
static int req_activate(p4g_server_s* v, SrvReq* r) {
    const long n = r->n;
    hipStream_t st = v->st;
    k_fill<double><<<1, 256, 0, st>>>(32, 0.0, r->S);
    k_fill<double><<<1, 256, 0, st>>>(1, 1.0, r->S + S_ONE);
    k_fill<double><<<1, 256, 0, st>>>(1, 1.0, r->S + 16 + S_ONE);
    const int gb = vblocks(n);
    k_fill<double><<<gb, 256, 0, st>>>(n, 0.0, r->x64);
    k_fill<float><<<gb, 256, 0, st>>>(n, 0.0f, r->pa);
    k_convert<double, float><<<gb, 256, 0, st>>>(n, r->b64, r->r32);
    k_dot2<double><<<gb, 256, 0, st>>>(n, r->b64, r->b64, r->part);
    k_finalize<<<1, FT, 0, st>>>(gb, r->part, r->S + 16, FIN_RAW, 0);
    LAUNCH_CHECK();
    r->pold = r->pa;
    r->pnew = r->pb;
    r->state = RQ_INIT;
    r->t_admit = wall_ms();
    return 0;
}

// ------------------------------------------------------------------ round

static int round_once(p4g_server_s* v) {
    Packer pk;
    std::vector<BLaunch> seq;
    // per-round CG phases (their item tables change every round)
    std::vector<DirItem> dir3, dir2;
    std::vector<FinItem> fa, frr, fraw, fbeta;
    std::vector<UpdateItem> up;
    std::vector<RepItem> rep3, rep2;
    long sd3 = 0, sd2 = 0, su = 0, sr3 = 0, sr2 = 0;
    for (SrvReq* r : v->active) {
        p4g_solver_s* s = r->sol;
        const LevelInfo& F = s->lv.back();
        r->replaced = false;
        if (r->state == RQ_ITER) {
            DirItem d;
            long& sd = r->dim == 3 ? sd3 : sd2;
            d.g = gref_stencil(F, sd, VecOf<float>::V);
            sd += gblocks(d.g);
            d.part = r->part;
            d.flag = nullptr;
            d.op = op32(s, F);
            d.src = SrcDir<float>{r->z32, r->pold, r->pnew, r->S + S_BCOEF};
            d.out = OutApplyDot<float>{r->w32};
            (r->dim == 3 ? dir3 : dir2).push_back(d);
            fa.push_back(FinItem{r->part, gblocks(d.g), r->S, FIN_ALPHA, 0, nullptr});
            std::swap(r->pold, r->pnew);   // pold: the current direction
            UpdateItem u;
            u.g.start = su;
            u.g.gx = vblocks(r->n);
            u.g.gy = u.g.gz = 1;
            su += u.g.gx;
            u.n = r->n;
            u.S = r->S;
            u.x = r->x64; u.r = r->r32; u.p = r->pold; u.w = r->w32;
            u.part = r->part;
            up.push_back(u);
            frr.push_back(FinItem{r->part, (long)u.g.gx, r->S, FIN_RR, 0, nullptr});
            if (r->replace_now) {
                RepItem q;
                long& sr = r->dim == 3 ? sr3 : sr2;
                q.g = gref_stencil(F, sr, VecOf<double>::V);
                sr += gblocks(q.g);
                q.part = r->part;
                q.flag = nullptr;
                q.op = op64(F);
                q.src = SrcArr<double>{r->x64};
                q.out = OutResidCvt<double, float>{r->b64, r->r32};
                (r->dim == 3 ? rep3 : rep2).push_back(q);
                fraw.push_back(FinItem{r->part, gblocks(q.g), r->S + 16, FIN_RAW, 0, nullptr});
                r->replace_now = false;
                r->replaced = true;
                r->nrep++;
            }
        }
        const LevelInfo& Ff = s->lv.back();
        fbeta.push_back(FinItem{r->part, nblocks(fused_grid(Ff.dim, Ff.mx, Ff.my, Ff.mz)), r->S,
                                FIN_BETA_PRECOND, r->state == RQ_INIT ? 1 : 0, nullptr});
    }
    auto add = [&](int kind, int dim, const void* data, size_t each, size_t cnt, long blocks) {
        if (!cnt) return;
        BLaunch a;
        a.kind = kind; a.dim = dim; a.nitems = (int)cnt; a.blocks = blocks;
        a.off = pk.put(data, each * cnt); a.len = each * cnt;
        seq.push_back(a);
    };
    add(BK_DIR, 3, dir3.data(), sizeof(DirItem), dir3.size(), sd3);
    add(BK_DIR, 2, dir2.data(), sizeof(DirItem), dir2.size(), sd2);
    add(BK_FIN, 0, fa.data(), sizeof(FinItem), fa.size(), (long)fa.size());
    add(BK_UPD, 0, up.data(), sizeof(UpdateItem), up.size(), su);
    add(BK_FIN, 0, frr.data(), sizeof(FinItem), frr.size(), (long)frr.size());
    add(BK_REP, 3, rep3.data(), sizeof(RepItem), rep3.size(), sr3);
    add(BK_REP, 2, rep2.data(), sizeof(RepItem), rep2.size(), sr2);
    add(BK_FIN, 0, fraw.data(), sizeof(FinItem), fraw.size(), (long)fraw.size());
    const size_t nbefore = seq.size();
    add(BK_FIN, 0, fbeta.data(), sizeof(FinItem), fbeta.size(), (long)fbeta.size());
    P4G_TRY(upload(v, pk, &v->d_round, &v->h_round, &v->round_cap, &v->round_bytes));
    for (size_t q = 0; q < nbefore; ++q) P4G_TRY(blaunch(v, seq[q], v->d_round, v->round_bytes));
    // V-cycle: launched directly while the active set changes, captured in a
    // hipGraph once the set has held for graph_after rounds (a capture costs
    // far more than one round's launches) and replayed until it changes
    std::vector<long> key;
    for (SrvReq* r : v->active) key.push_back(r->id);
    if (key != v->vc_key) {
        v->vc_key = key;
        v->vc_stable = 0;
        v->vc_built = false;
    } else {
        v->vc_stable++;
    }
    const bool want_graph = v->use_graph && v->vc_stable >= v->graph_after;
    if (!v->vc_built || (want_graph && (key != v->graph_key || !v->gexec))) {
        Packer vp;
        v->vseq.clear();   // build_vcycle appends; the old entries index the old table
        v->vc_built = false;
        P4G_TRY(build_vcycle(v, vp, v->vseq));
        P4G_TRY(upload(v, vp, &v->d_vc, &v->h_vc, &v->vc_cap, &v->vc_bytes));
        v->vc_built = true;
        if (v->gexec) { hipGraphExecDestroy(v->gexec); v->gexec = nullptr; }
        if (v->graph) { hipGraphDestroy(v->graph); v->graph = nullptr; }
        v->graph_key.clear();
        const std::vector<BLaunch>& vseq = v->vseq;
        if (want_graph) {
            HIP_TRY(hipStreamSynchronize(v->st));
            HIP_TRY(hipStreamBeginCapture(v->st, hipStreamCaptureModeThreadLocal));
            int e = 0;
            for (const BLaunch& L : vseq)
                if (!e) e = blaunch(v, L, v->d_vc, v->vc_bytes);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(v->st, &g);
            if (e) { if (g) hipGraphDestroy(g); return e; }
            if (ce != hipSuccess) return set_error(PETSC_ERR_LIB, "serving: graph capture: %s", hipGetErrorString(ce));
            v->graph = g;
            HIP_TRY(hipGraphInstantiate(&v->gexec, g, nullptr, nullptr, 0));
            v->graph_key = key;
            v->captures++;
        }
    }
    if (want_graph && v->gexec) {
        HIP_TRY(hipGraphLaunch(v->gexec, v->st));
    } else {
        for (const BLaunch& L : v->vseq) P4G_TRY(blaunch(v, L, v->d_vc, v->vc_bytes));
    }
    for (size_t q = nbefore; q < seq.size(); ++q) P4G_TRY(blaunch(v, seq[q], v->d_round, v->round_bytes));
    HIP_TRY(hipMemcpyAsync(v->hS, v->S, sizeof(double) * 32 * v->max_slots, hipMemcpyDeviceToHost, v->st));
    HIP_TRY(hipStreamSynchronize(v->st));
    v->rounds++;
    return 0;
}

// host decisions after a round; returns true when the request finished
static bool req_step(p4g_server_s* v, SrvReq* r) {
    p4g_solver_s* s = r->sol;
    const double* hs = v->hS + 32 * r->slot;
    const double* ho = hs + 16;
    if (r->state == RQ_INIT) {
        r->bnorm = std::sqrt(ho[S_TMP0]);
        r->target = std::max(s->rtol * r->bnorm, s->atol);
        r->rr_last = r->bnorm * r->bnorm;
        r->rn = r->bnorm;
        r->rz_old = hs[S_RZ];
        r->state = RQ_ITER;
        if (r->bnorm <= s->atol) { r->reason = P4G_CONVERGED_ATOL; return true; }
        return false;
    }
    r->its++;
    const double dpi = hs[S_PW];
    if (dpi == 0.0 || (r->its > 1 && ((dpi > 0) != (r->dpi_old > 0)))) { r->reason = P4G_DIVERGED_INDEFINITE_MAT; return true; }
    r->dpi_old = dpi;
    double rn;
    if (r->replaced) { rn = std::sqrt(ho[S_TMP0]); r->rr_last = rn * rn; }
    else rn = std::sqrt(hs[S_RR]);
    r->rn = rn;
    if (std::isnan(rn) || std::isinf(rn)) { r->reason = P4G_DIVERGED_NANORINF; return true; }
    if (r->replaced && rn <= r->target) { r->reason = rn < s->atol ? P4G_CONVERGED_ATOL : P4G_CONVERGED_RTOL; return true; }
    if (rn >= s->dtol * r->bnorm) { r->reason = P4G_DIVERGED_DTOL; return true; }
    if (r->its >= s->maxit) { r->reason = P4G_DIVERGED_ITS; return true; }
    if (hs[S_RZ] * r->rz_old < 0.0) { r->its++; r->reason = P4G_DIVERGED_INDEFINITE_PC; return true; }   // beta changed sign
    r->rz_old = hs[S_RZ];
    if (hs[S_RZ] == 0.0 || (!r->replaced && (rn <= r->target || rn * rn <= s->reliable_delta * s->reliable_delta * r->rr_last)))
        r->replace_now = true;
    return false;
}

// the result's copy into the pinned staging buffer, on the copy stream after
// the request's last kernel (the rounds do not wait for it; nothing of this
// request's is reused until the completion thread has seen the copy's event)
static int req_finish(p4g_server_s* v, SrvReq* r) {
    HIP_TRY(hipEventRecord(r->ev_k, v->st));   // the request's last kernel
    HIP_TRY(hipStreamWaitEvent(v->cst, r->ev_k, 0));
    HIP_TRY(hipMemcpyAsync(r->pin, r->x64, r->n * 8, hipMemcpyDeviceToHost, v->cst));
    HIP_TRY(hipEventRecord(r->ev, v->cst));
    return 0;
}

// completion thread: waits for each result copy in order, copies the pinned
// result into the caller's array, then releases the request's device vectors
// to the pool and its staging buffer, and reports it done
static void copy_loop(p4g_server_s* v) {
    hipSetDevice(v->device);
    for (;;) {
        SrvReq* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(v->mu);
            v->cv_copy.wait(lk, [&] { return v->stop_copy || !v->copying.empty(); });
            if (v->copying.empty()) break;   // stop_copy and nothing left
            r = v->copying.front();
            v->copying.pop_front();
        }
        const hipError_t he = hipEventSynchronize(r->ev);
        if (he == hipSuccess) std::memcpy(r->hx, r->pin, r->n * sizeof(double));
        const double t = wall_ms();
        std::lock_guard<std::mutex> lk(v->mu);
        hipEventDestroy(r->ev);
        hipEventDestroy(r->ev_in);
        hipEventDestroy(r->ev_k);
        r->ev = r->ev_in = r->ev_k = nullptr;
        v->ppool[r->n].push_back(r->pin);
        r->pin = nullptr;
        v->dpool[r->layout].push_back(std::move(r->allocs));
        r->allocs.clear();
        v->pool_bytes += r->bytes;
        v->used -= r->bytes;
        if (he != hipSuccess) {
            r->state = RQ_FAILED;
            r->err = hipGetErrorString(he);
        } else {
            r->t_done = t;
            r->state = RQ_DONE;
            v->lat_ms.push_back(r->t_done - r->t_submit);
            v->queue_ms.push_back(r->t_admit - r->t_submit);
            v->completed++;
            v->dof_done += (double)r->n;
            v->t_last = r->t_done;
        }
        v->cv_done.notify_all();
        v->cv.notify_all();   // the pool may now hold what a queued request needs
    }
}

static void serve_loop(p4g_server_s* v) {
    hipSetDevice(v->device);
    for (;;) {
        std::vector<SrvReq*> newly;
        {
            std::unique_lock<std::mutex> lk(v->mu);
            v->cv.wait(lk, [&] {
                return v->stop || !v->queue.empty() || !v->active.empty() || !v->loading.empty();
            });
            if (v->stop && v->active.empty() && v->queue.empty() && v->loading.empty()) break;
            // admission in arrival order, skipping requests that do not fit yet
            for (auto it = v->queue.begin(); it != v->queue.end();) {
                SrvReq* r = *it;
                if ((int)(v->active.size() + v->loading.size() + newly.size()) >= v->max_slots) break;
                if (v->used + r->bytes <= v->budget) {
                    v->used += r->bytes;
                    v->peak = std::max(v->peak, v->used);
                    newly.push_back(r);
                    it = v->queue.erase(it);
                } else {
                    ++it;
                }
            }
        }
        int e = 0;
        for (SrvReq* r : newly) {   // vectors, slot and b's upload on the copy stream
            v->loading.push_back(r);
            if (!e) e = req_admit(v, r);
        }
        // requests whose upload has finished join the rounds; with none
        // active, wait for the oldest upload instead of spinning
        if (!e && !v->loading.empty() && v->active.empty()) {
            const hipError_t he = hipEventSynchronize(v->loading.front()->ev_in);
            if (he != hipSuccess) e = set_error(PETSC_ERR_LIB, "serving: upload: %s", hipGetErrorString(he));
        }
        for (auto it = v->loading.begin(); !e && it != v->loading.end();) {
            const hipError_t q = hipEventQuery((*it)->ev_in);
            if (q == hipSuccess) {
                e = req_activate(v, *it);
                std::lock_guard<std::mutex> lk(v->mu);
                v->active.push_back(*it);
                it = v->loading.erase(it);
            } else if (q == hipErrorNotReady) {
                ++it;
            } else {
                e = set_error(PETSC_ERR_LIB, "serving: upload: %s", hipGetErrorString(q));
            }
        }
        if (!e && v->active.empty()) continue;
        if (!e) e = round_once(v);
        std::vector<SrvReq*> done;
        if (!e)
            for (SrvReq* r : v->active)
                if (req_step(v, r)) done.push_back(r);
        for (SrvReq* r : done)
            if (!e) e = req_finish(v, r);
        std::lock_guard<std::mutex> lk(v->mu);
        if (e) {
            v->fatal = p4g_last_error();
            auto drop = [&](SrvReq* r, bool held) {
                r->state = RQ_FAILED;
                r->err = v->fatal;
                if (held) req_free(r);
                if (r->pin) { v->ppool[r->n].push_back(r->pin); r->pin = nullptr; }
                for (hipEvent_t* q : {&r->ev, &r->ev_in, &r->ev_k})
                    if (*q) { hipEventDestroy(*q); *q = nullptr; }
            };
            for (SrvReq* r : v->active) drop(r, true);
            for (SrvReq* r : v->loading) drop(r, true);
            for (SrvReq* r : v->queue) drop(r, false);
            v->active.clear();
            v->loading.clear();
            v->queue.clear();
            size_t held = 0;   // results still being copied keep their vectors until the completion thread is done
            for (SrvReq* r : v->copying) held += r->bytes;
            v->used = held;
            v->cv_done.notify_all();
            continue;
        }
        for (SrvReq* r : done) {   // to the completion thread (budget, pools and statistics there)
            v->free_slots.push_back(r->slot);
            v->active.erase(std::find(v->active.begin(), v->active.end(), r));
            v->copying.push_back(r);
        }
        if (!done.empty()) v->cv_copy.notify_all();
    }
}

static double percentile(std::vector<double> x, double q) {
    if (x.empty()) return 0.0;
    std::sort(x.begin(), x.end());
    const double k = (x.size() - 1) * q;
    const size_t lo = (size_t)std::floor(k), hi = (size_t)std::ceil(k);
    return x[lo] + (x[hi] - x[lo]) * (k - lo);
}

}  // namespace p4g

// ================================================================== C ABI

extern "C" {

int p4g_server_create(size_t hbm_budget_bytes, const char* options, p4g_Server* out) {
    clear_error();
    if (!out) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_create: null output");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(PETSC_ERR_LIB, "p4g_server_create: no HIP device is available");
    Options o;
    o.parse(options);
    p4g_server_s* v = new p4g_server_s;
    v->budget = hbm_budget_bytes;
    v->use_graph = o.getb("p4g_serve_graph", true);
    v->graph_after = std::max(0, o.geti("p4g_serve_graph_after", 32));
    v->max_slots = std::max(1, std::min(1024, o.geti("p4g_serve_max_slots", 64)));
    hipGetDevice(&v->device);
    int e = 0;
    // non-blocking (default priority): while the worker captures a graph on
    // it, synchronous copies on the legacy stream from other threads do not
    // depend on it (the server orders its own work with explicit syncs)
    if (hipStreamCreateWithFlags(&v->st, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&v->cst, hipStreamNonBlocking) != hipSuccess)
        e = set_error(PETSC_ERR_LIB, "p4g_server_create: stream");
    if (!e && hipMalloc(&v->S, sizeof(double) * 32 * v->max_slots) != hipSuccess)
        e = set_error(PETSC_ERR_MEM, "p4g_server_create: scalar blocks");
    if (!e && hipHostMalloc((void**)&v->hS, sizeof(double) * 32 * v->max_slots, hipHostMallocDefault) != hipSuccess)
        e = set_error(PETSC_ERR_MEM, "p4g_server_create: pinned scalars");
    const long staging_mb = std::max(0, o.geti("p4g_serve_staging_mb", 1024));
    if (!e && staging_mb > 0) {
        v->arena_size = (size_t)staging_mb << 20;
        if (hipHostMalloc((void**)&v->arena, v->arena_size, hipHostMallocDefault) != hipSuccess)
            e = set_error(PETSC_ERR_MEM, "p4g_server_create: %ld MiB pinned staging", staging_mb);
        else
            v->arena_free[0] = v->arena_size;
    }
    if (e) {
        if (v->S) hipFree(v->S);
        if (v->hS) hipHostFree(v->hS);
        if (v->arena) hipHostFree(v->arena);
        if (v->st) hipStreamDestroy(v->st);
        if (v->cst) hipStreamDestroy(v->cst);
        delete v;
        return e;
    }
    for (int q = v->max_slots - 1; q >= 0; --q) v->free_slots.push_back(q);
    v->th = std::thread(serve_loop, v);
    for (int q = 0; q < 4; ++q) v->th_copy.emplace_back(copy_loop, v);   // results of several requests at once
    *out = v;
    return 0;
}

int p4g_server_request_bytes(p4g_Solver s, size_t* bytes) {
    clear_error();
    if (!s || !bytes) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_request_bytes: null argument");
    P4G_TRY(solver_setup(s));
    *bytes = req_bytes(s);
    return 0;
}

int p4g_server_submit(p4g_Server v, p4g_Solver s, const double* b, double* x, long* id) {
    clear_error();
    if (!v || !s || !b || !x || !id) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_submit: null argument");
    P4G_TRY(solver_setup(s));
    P4G_TRY(serve_supported(s));
    SrvReq* r = new SrvReq;
    r->sol = s;
    r->hb = b;
    r->hx = x;
    r->bytes = req_bytes(s);
    r->n = s->lv.back().n;
    r->t_submit = wall_ms();
    // b into pinned staging here, in the caller's thread; the worker's
    // upload is then an asynchronous pinned copy
    if (r->bytes <= v->budget) {
        int pe = pin_get(v, r->n, &r->pin);
        for (hipEvent_t* q : {&r->ev, &r->ev_in, &r->ev_k})
            if (!pe && hipEventCreateWithFlags(q, hipEventDisableTiming) != hipSuccess)
                pe = set_error(PETSC_ERR_LIB, "p4g_server_submit: event");
        if (pe) {
            if (r->pin) { std::lock_guard<std::mutex> lk(v->mu); v->ppool[r->n].push_back(r->pin); }
            for (hipEvent_t* q : {&r->ev, &r->ev_in, &r->ev_k})
                if (*q) hipEventDestroy(*q);
            delete r;
            return pe;
        }
        std::memcpy(r->pin, b, r->n * sizeof(double));
    }
    std::lock_guard<std::mutex> lk(v->mu);
    if (!v->fatal.empty()) {
        if (r->pin) v->ppool[r->n].push_back(r->pin);
        for (hipEvent_t* q : {&r->ev, &r->ev_in, &r->ev_k})
            if (*q) hipEventDestroy(*q);
        delete r;
        return set_error(PETSC_ERR_LIB, "server failed: %s", v->fatal.c_str());
    }
    r->id = v->next_id++;
    *id = r->id;
    v->all[r->id] = r;
    v->submitted++;
    if (v->t_first < 0) v->t_first = r->t_submit;
    if (r->bytes > v->budget) {
        r->state = RQ_FAILED;
        r->err = "request footprint exceeds the server's HBM budget";
        v->rejected++;
        return set_error(PETSC_ERR_MEM, "p4g_server_submit: request needs %zu bytes, budget %zu", r->bytes, v->budget);
    }
    v->queue.push_back(r);
    v->cv.notify_all();
    return 0;
}

int p4g_server_poll(p4g_Server v, long id, int* state) {
    if (!v || !state) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_poll: null argument");
    std::lock_guard<std::mutex> lk(v->mu);
    auto it = v->all.find(id);
    if (it == v->all.end()) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_server_poll: unknown request %ld", id);
    *state = it->second->state;
    return 0;
}

int p4g_server_wait(p4g_Server v, long id, p4g_Report* rep) {
    clear_error();
    if (!v) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_wait: null server");
    std::unique_lock<std::mutex> lk(v->mu);
    auto it = v->all.find(id);
    if (it == v->all.end()) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_server_wait: unknown request %ld", id);
    SrvReq* r = it->second;
    v->cv_done.wait(lk, [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; });
    if (r->state == RQ_FAILED) return set_error(PETSC_ERR_LIB, "request %ld failed: %s", id, r->err.c_str());
    if (rep) {
        std::memset(rep, 0, sizeof *rep);
        rep->its = r->its;
        rep->reason = r->reason;
        rep->inner_its = r->nrep;
        rep->nlevels = r->nlev;
        rep->rnorm0 = r->bnorm;
        rep->rnorm = r->rn;
        rep->true_relres = r->bnorm > 0 ? r->rn / r->bnorm : 0.0;
        rep->setup_ms = r->t_admit - r->t_submit;   // queueing
        rep->solve_ms = r->t_done - r->t_submit;    // latency: submit -> result on host
    }
    return 0;
}

int p4g_server_stats(p4g_Server v, p4g_ServerStats* st) {
    if (!v || !st) return set_error(PETSC_ERR_ARG_WRONG, "p4g_server_stats: null argument");
    std::lock_guard<std::mutex> lk(v->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = v->submitted;
    st->completed = v->completed;
    st->rejected = v->rejected;
    st->active = (long)v->active.size();
    st->queued = (long)v->queue.size();
    st->rounds = v->rounds;
    st->graph_captures = v->captures;
    st->p50_ms = percentile(v->lat_ms, 0.5);
    st->p99_ms = percentile(v->lat_ms, 0.99);
    st->max_ms = v->lat_ms.empty() ? 0.0 : *std::max_element(v->lat_ms.begin(), v->lat_ms.end());
    double sum = 0.0;
    for (double q : v->lat_ms) sum += q;
    st->mean_ms = v->lat_ms.empty() ? 0.0 : sum / v->lat_ms.size();
    st->queue_p50_ms = percentile(v->queue_ms, 0.5);
    st->wall_s = v->t_first >= 0 ? (v->t_last - v->t_first) * 1e-3 : 0.0;
    st->dof_per_s = st->wall_s > 0 ? v->dof_done / st->wall_s : 0.0;
    st->peak_bytes = (double)v->peak;
    return 0;
}

int p4g_server_destroy(p4g_Server* pv) {
    if (!pv || !*pv) return 0;
    p4g_server_s* v = *pv;
    {
        std::lock_guard<std::mutex> lk(v->mu);
        v->stop = true;
        v->cv.notify_all();
    }
    if (v->th.joinable()) v->th.join();
    {
        std::lock_guard<std::mutex> lk(v->mu);
        v->stop_copy = true;
        v->cv_copy.notify_all();
    }
    for (std::thread& t : v->th_copy)
        if (t.joinable()) t.join();
    for (auto& kv : v->all) {
        for (hipEvent_t q : {kv.second->ev, kv.second->ev_in, kv.second->ev_k})
            if (q) hipEventDestroy(q);
        if (kv.second->pin) hipHostFree(kv.second->pin);
        delete kv.second;
    }
    for (auto& kv : v->dpool)
        for (auto& set : kv.second)
            for (void* p : set) hipFree(p);
    for (auto& kv : v->ppool)
        for (double* p : kv.second)
            if (!(v->arena && (char*)p >= v->arena && (char*)p < v->arena + v->arena_size)) hipHostFree(p);
    if (v->arena) hipHostFree(v->arena);
    if (v->gexec) hipGraphExecDestroy(v->gexec);
    if (v->graph) hipGraphDestroy(v->graph);
    if (v->d_round) hipFree(v->d_round);
    if (v->h_round) hipHostFree(v->h_round);
    if (v->d_vc) hipFree(v->d_vc);
    if (v->h_vc) hipHostFree(v->h_vc);
    if (v->S) hipFree(v->S);
    if (v->hS) hipHostFree(v->hS);
    if (v->st) hipStreamDestroy(v->st);
    if (v->cst) hipStreamDestroy(v->cst);
    delete v;
    *pv = nullptr;
    return 0;
}

}  // extern "C"
