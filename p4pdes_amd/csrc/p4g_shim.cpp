// p4g_shim.cpp -- the PETSc-shaped host surface (p4g_petsc.h): grid queries,
// the recording matrix, host vectors, errors and options.  Set-up only;
// nothing here runs inside a solve.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "p4g_internal.hpp"

namespace p4g {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

// ----------------------------------------------------------------- options

static bool is_value(const char* t) {
    if (t[0] != '-') return true;
    return (t[1] >= '0' && t[1] <= '9') || t[1] == '.';
}

void Options::parse(const char* s) {
    keys.clear();
    vals.clear();
    if (!s) return;
    std::vector<std::string> tok;
    const char* p = s;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\n') ++p;
        if (!*p) break;
        const char* q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n') ++q;
        tok.emplace_back(p, q - p);
        p = q;
    }
    for (size_t a = 0; a < tok.size(); ++a) {
        const std::string& t = tok[a];
        if (t.size() < 2 || t[0] != '-' || is_value(t.c_str())) continue;
        keys.push_back(t.substr(1));
        if (a + 1 < tok.size() && is_value(tok[a + 1].c_str())) vals.push_back(tok[++a]);
        else vals.push_back("");
    }
}

const std::string* Options::get(const std::string& k) const {
    for (size_t q = keys.size(); q-- > 0;)
        if (keys[q] == k) return &vals[q];
    return nullptr;
}
bool Options::has(const std::string& k) const { return get(k) != nullptr; }
int Options::geti(const std::string& k, int def) const {
    const std::string* v = get(k);
    return (v && !v->empty()) ? atoi(v->c_str()) : def;
}
double Options::getr(const std::string& k, double def) const {
    const std::string* v = get(k);
    return (v && !v->empty()) ? atof(v->c_str()) : def;
}
std::string Options::gets(const std::string& k, const std::string& def) const {
    const std::string* v = get(k);
    return v ? *v : def;
}
bool Options::getb(const std::string& k, bool def) const {
    const std::string* v = get(k);
    if (!v) return def;
    if (v->empty()) return true;
    return !(*v == "false" || *v == "0" || *v == "no");
}

// ------------------------------------------------------------------- views

void info_of(p4g_da_s* da, DMDALocalInfo* info) {
    std::memset(info, 0, sizeof *info);
    info->da = da;
    info->dim = da->dim;
    info->dof = 1;
    info->sw = 1;
    info->mx = da->mx;
    info->my = da->dim > 1 ? da->my : 1;
    info->mz = da->dim > 2 ? da->mz : 1;
    info->xm = info->mx; info->ym = info->my; info->zm = info->mz;
    info->gxm = info->mx; info->gym = info->my; info->gzm = info->mz;
}

void* make_view(p4g_da_s* da, double* base, std::vector<double*>& rows,
                std::vector<double**>& planes) {
    if (da->dim == 1) return base;
    const int my = da->my, mz = da->dim > 2 ? da->mz : 1;
    rows.resize((size_t)my * mz);
    for (long r = 0; r < (long)my * mz; ++r) rows[r] = base + r * (long)da->mx;
    if (da->dim == 2) return rows.data();
    planes.resize(mz);
    for (int k = 0; k < mz; ++k) planes[k] = rows.data() + (long)k * my;
    return planes.data();
}

// big-endian writer / reader of PETSc binary files
const int32_t kVecClassId = 1211214, kMatClassId = 1211216;

BinaryFile::~BinaryFile() {
    if (fp) fclose(fp);
}
bool BinaryFile::open(const char* path, const char* mode) {
    fp = fopen(path, mode);
    return fp != nullptr;
}
bool BinaryFile::close() {
    bool ok = flush();
    if (fp && fclose(fp) != 0) ok = false;
    fp = nullptr;
    return ok && !failed;
}
bool BinaryFile::flush() {
    if (!fp || buf.empty()) return true;
    if (fwrite(buf.data(), 1, buf.size(), fp) != buf.size()) failed = true;
    buf.clear();
    return !failed;
}
void BinaryFile::bytes(const void* p, size_t k) {
    const unsigned char* c = (const unsigned char*)p;
    buf.insert(buf.end(), c, c + k);
    if (buf.size() >= (1u << 20)) flush();
}
void BinaryFile::i32(int32_t v) {
    const uint32_t u = __builtin_bswap32((uint32_t)v);
    bytes(&u, 4);
}
void BinaryFile::f64(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    u = __builtin_bswap64(u);
    bytes(&u, 8);
}
void BinaryFile::f64s(const double* x, long n) {
    for (long q = 0; q < n; ++q) f64(x[q]);
}
bool BinaryFile::read_i32(int32_t& v) {
    uint32_t u;
    if (fread(&u, 4, 1, fp) != 1) return false;
    v = (int32_t)__builtin_bswap32(u);
    return true;
}
bool BinaryFile::read_f64s(double* x, long n) {
    std::vector<uint64_t> tmp(1 << 16);
    for (long q = 0; q < n;) {
        const long k = std::min<long>((long)tmp.size(), n - q);
        if (fread(tmp.data(), 8, (size_t)k, fp) != (size_t)k) return false;
        for (long e = 0; e < k; ++e) {
            const uint64_t u = __builtin_bswap64(tmp[e]);
            std::memcpy(&x[q + e], &u, 8);
        }
        q += k;
    }
    return true;
}
bool BinaryFile::skip(long k) { return fseek(fp, k, SEEK_CUR) == 0; }

}  // namespace p4g

using namespace p4g;

extern "C" {

PetscErrorCode p4g_seterr(PetscErrorCode code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return set_error(code, "%s", buf);
}

const char* p4g_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------ PETSc binary vectors

int p4g_vec_view_binary(const char* path, long n, const double* x, int append) {
    clear_error();
    if (!path || n < 0 || (n > 0 && !x)) return set_error(PETSC_ERR_ARG_WRONG, "p4g_vec_view_binary: bad argument");
    if (n > 2147483647L) return set_error(PETSC_ERR_ARG_OUTOFRANGE, "p4g_vec_view_binary: length exceeds int32");
    BinaryFile f;
    if (!f.open(path, append ? "ab" : "wb")) return set_error(PETSC_ERR_FILE_OPEN, "cannot open %s", path);
    f.i32(kVecClassId);
    f.i32((int32_t)n);
    f.f64s(x, n);
    if (!f.close()) return set_error(PETSC_ERR_FILE_WRITE, "write to %s failed", path);
    return 0;
}

int p4g_vec_load_binary(const char* path, int index, long* n, double* x) {
    clear_error();
    if (!path || index < 0 || !n) return set_error(PETSC_ERR_ARG_WRONG, "p4g_vec_load_binary: bad argument");
    BinaryFile f;
    if (!f.open(path, "rb")) return set_error(PETSC_ERR_FILE_OPEN, "cannot open %s", path);
    for (int k = 0;; ++k) {
        int32_t id = 0, len = 0;
        if (!f.read_i32(id)) return set_error(PETSC_ERR_FILE_READ, "%s: object %d not found", path, index);
        if (id != kVecClassId) {
            if (k == index) return set_error(PETSC_ERR_FILE_UNEXPECTED, "%s: object %d is not a PETSc Vec", path, index);
            // skip a matrix: sizes, row lengths, column indices, values
            if (id != kMatClassId) return set_error(PETSC_ERR_FILE_UNEXPECTED, "%s: unknown class id %d", path, id);
            int32_t m = 0, nc = 0, nz = 0;
            if (!f.read_i32(m) || !f.read_i32(nc) || !f.read_i32(nz) || m < 0 || nz < 0 ||
                !f.skip(4L * m + 4L * nz + 8L * nz))
                return set_error(PETSC_ERR_FILE_READ, "%s: truncated matrix", path);
            continue;
        }
        if (!f.read_i32(len) || len < 0) return set_error(PETSC_ERR_FILE_READ, "%s: truncated vector header", path);
        if (k < index) {
            if (!f.skip(8L * len)) return set_error(PETSC_ERR_FILE_READ, "%s: truncated vector", path);
            continue;
        }
        *n = len;
        if (x && !f.read_f64s(x, len)) return set_error(PETSC_ERR_FILE_READ, "%s: truncated vector", path);
        return 0;
    }
}

// ---------------------------------------------------------------- DM queries

PetscErrorCode DMGetBoundingBox(DM da, PetscReal gmin[], PetscReal gmax[]) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "DMGetBoundingBox: null DM");
    for (int d = 0; d < da->dim; ++d) { gmin[d] = da->lo[d]; gmax[d] = da->lo[d] + da->L[d]; }
    return 0;
}

PetscErrorCode DMDASetUniformCoordinates(DM da, PetscReal xmin, PetscReal xmax, PetscReal ymin, PetscReal ymax,
                                         PetscReal zmin, PetscReal zmax) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "DMDASetUniformCoordinates: null DM");
    const double a[3] = {xmin, ymin, zmin}, b[3] = {xmax, ymax, zmax};
    for (int d = 0; d < da->dim; ++d) {
        if (!(b[d] > a[d]))
            return set_error(PETSC_ERR_ARG_OUTOFRANGE, "DMDASetUniformCoordinates: empty interval in dimension %d", d);
        da->lo[d] = a[d];
        da->L[d] = b[d] - a[d];
    }
    return 0;
}

PetscErrorCode DMDAGetLocalInfo(DM da, DMDALocalInfo* info) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "DMDAGetLocalInfo: null DM");
    info_of(da, info);
    return 0;
}

PetscErrorCode DMSetApplicationContext(DM da, void* ctx) {
    da->appctx = ctx;
    return 0;
}

PetscErrorCode DMGetApplicationContext(DM da, void* ctx) {
    *(void**)ctx = da->appctx;
    return 0;
}

// --------------------------------------------------------- recording matrix

static int mismatch(Mat A, const char* what, int i, int j, int k) {
    if (A->bad++ == 0) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s at row (k,j,i)=(%d,%d,%d)", what, k, j, i);
        A->why = buf;
    }
    return 0;
}

static bool same(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fabs(b); }

PetscErrorCode MatSetValuesStencil(Mat A, PetscInt m, const MatStencil idxm[], PetscInt n,
                                   const MatStencil idxn[], const PetscScalar v[],
                                   InsertMode mode) {
    if (!A || !A->da) return set_error(PETSC_ERR_ARG_WRONG, "MatSetValuesStencil: null Mat");
    if (mode != INSERT_VALUES)
        return set_error(PETSC_ERR_SUP, "MatSetValuesStencil: only INSERT_VALUES is supported");
    p4g_da_s* da = A->da;
    const int dim = da->dim, mx = da->mx, my = dim > 1 ? da->my : 1, mz = dim > 2 ? da->mz : 1;
    for (PetscInt r = 0; r < m; ++r) {
        const int i = idxm[r].i, j = dim > 1 ? idxm[r].j : 0, k = dim > 2 ? idxm[r].k : 0;
        if (i < 0 || i >= mx || j < 0 || j >= my || k < 0 || k >= mz)
            return set_error(PETSC_ERR_ARG_OUTOFRANGE, "MatSetValuesStencil: row out of range");
        const long row = ((long)k * my + j) * mx + i;
        if (!A->seen[row]) { A->seen[row] = 1; A->rows++; }
        const bool bnd = i == 0 || i == mx - 1 || (dim > 1 && (j == 0 || j == my - 1)) ||
                         (dim > 2 && (k == 0 || k == mz - 1));
        // couplings expected toward interior neighbours only
        int expect = 0;
        if (!bnd) {
            expect += (i - 1 > 0) + (i + 1 < mx - 1);
            if (dim > 1) expect += (j - 1 > 0) + (j + 1 < my - 1);
            if (dim > 2) expect += (k - 1 > 0) + (k + 1 < mz - 1);
        }
        int got = 0;
        bool hasdiag = false;
        for (PetscInt c = 0; c < n; ++c) {
            const int di = idxn[c].i - i, dj = dim > 1 ? idxn[c].j - j : 0,
                      dk = dim > 2 ? idxn[c].k - k : 0;
            const double val = v[r * n + c];
            if (A->keep) {
                const int ci = idxn[c].i, cj = dim > 1 ? idxn[c].j : 0, ck = dim > 2 ? idxn[c].k : 0;
                if (ci < 0 || ci >= mx || cj < 0 || cj >= my || ck < 0 || ck >= mz)
                    return set_error(PETSC_ERR_ARG_OUTOFRANGE, "MatSetValuesStencil: column out of range");
                A->er.push_back(row);
                A->ec.push_back(((long)ck * my + cj) * mx + ci);
                A->ev.push_back(val);
            }
            if (di == 0 && dj == 0 && dk == 0) {
                hasdiag = true;
                if (bnd) {
                    if (!A->have_b) { A->have_b = true; A->db = val; }
                    else if (!same(val, A->db)) mismatch(A, "boundary diagonal differs", i, j, k);
                } else {
                    if (!A->have_i) { A->have_i = true; A->d = val; }
                    else if (!same(val, A->d)) mismatch(A, "interior diagonal differs", i, j, k);
                }
                continue;
            }
            const int dir = (std::abs(di) == 1 && dj == 0 && dk == 0) ? 0
                          : (di == 0 && std::abs(dj) == 1 && dk == 0) ? 1
                          : (di == 0 && dj == 0 && std::abs(dk) == 1) ? 2 : -1;
            if (bnd || dir < 0) {
                if (val != 0.0) mismatch(A, "entry outside the star-stencil pattern", i, j, k);
                continue;
            }
            const int ni = i + di, nj = j + dj, nk = k + dk;
            const bool nbnd = ni == 0 || ni == mx - 1 || (dim > 1 && (nj == 0 || nj == my - 1)) ||
                              (dim > 2 && (nk == 0 || nk == mz - 1));
            if (nbnd) {
                if (val != 0.0) mismatch(A, "coupling to a boundary vertex", i, j, k);
                continue;
            }
            ++got;
            if (!A->have_s[dir]) { A->have_s[dir] = true; A->s[dir] = -val; }
            else if (!same(-val, A->s[dir])) mismatch(A, "coupling differs", i, j, k);
        }
        if (!hasdiag) mismatch(A, "missing diagonal", i, j, k);
        if (got != expect) mismatch(A, "missing coupling", i, j, k);
    }
    return 0;
}

PetscErrorCode MatAssemblyBegin(Mat, MatAssemblyType) { return 0; }
PetscErrorCode MatAssemblyEnd(Mat, MatAssemblyType) { return 0; }

// ------------------------------------------------------------------ vectors

PetscErrorCode DMCreateGlobalVector(DM da, Vec* v) {
    if (!da) return set_error(PETSC_ERR_ARG_WRONG, "DMCreateGlobalVector: null DM");
    p4g_vec_s* x = new p4g_vec_s;
    x->da = da;
    x->a.assign(da->n(), 0.0);
    *v = x;
    return 0;
}
PetscErrorCode DMGetGlobalVector(DM da, Vec* v) { return DMCreateGlobalVector(da, v); }
PetscErrorCode DMRestoreGlobalVector(DM, Vec* v) { return VecDestroy(v); }
PetscErrorCode VecDestroy(Vec* v) {
    if (v && *v) { delete *v; *v = nullptr; }
    return 0;
}
PetscErrorCode VecSet(Vec v, PetscScalar a) {
    for (double& q : v->a) q = a;
    return 0;
}
PetscErrorCode VecAXPY(Vec y, PetscScalar a, Vec x) {
    if (y->a.size() != x->a.size()) return set_error(PETSC_ERR_ARG_SIZ, "VecAXPY: sizes differ");
    for (size_t q = 0; q < y->a.size(); ++q) y->a[q] += a * x->a[q];
    return 0;
}
PetscErrorCode VecNorm(Vec v, NormType t, PetscReal* nrm) {
    double s = 0.0;
    if (t == NORM_INFINITY) {
        for (double q : v->a) s = std::fabs(q) > s ? std::fabs(q) : s;
    } else if (t == NORM_1) {
        for (double q : v->a) s += std::fabs(q);
    } else {
        for (double q : v->a) s += q * q;
        s = std::sqrt(s);
    }
    *nrm = s;
    return 0;
}
PetscErrorCode VecGetSize(Vec v, PetscInt* n) {
    *n = (PetscInt)v->a.size();
    return 0;
}
PetscErrorCode VecGetArray(Vec v, PetscScalar** a) {
    *a = v->a.data();
    return 0;
}
PetscErrorCode VecRestoreArray(Vec, PetscScalar**) { return 0; }

PetscErrorCode PetscRandomCreate(MPI_Comm, PetscRandom* r) {
    *r = new p4g_random_s;
    return 0;
}
PetscErrorCode PetscRandomDestroy(PetscRandom* r) {
    if (r && *r) { delete *r; *r = nullptr; }
    return 0;
}
// uniform [0,1) from a splitmix64 stream
PetscErrorCode VecSetRandom(Vec v, PetscRandom r) {
    for (double& q : v->a) {
        uint64_t z = r->seed + 0x9E3779B97F4A7C15ull * (++r->counter);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        q = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    return 0;
}

PetscErrorCode DMDAVecGetArray(DM da, Vec v, void* array) {
    if (!da || !v || v->a.size() != (size_t)da->n())
        return set_error(PETSC_ERR_ARG_WRONG, "DMDAVecGetArray: vector does not match DM");
    *(void**)array = make_view(da, v->a.data(), v->rows, v->planes);
    return 0;
}
PetscErrorCode DMDAVecRestoreArray(DM, Vec, void* array) {
    *(void**)array = nullptr;
    return 0;
}

static thread_local double g_flops = 0.0;
PetscErrorCode PetscLogFlops(PetscReal n) {
    g_flops += n;
    return 0;
}

PetscErrorCode PetscPrintf(MPI_Comm, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    return 0;
}

}  // extern "C"
