// p4g_internal.hpp -- host-side types shared by the p4g library sources.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "p4g.h"

// ------------------------------------------------------------------ errors

namespace p4g {
int set_error(int code, const char* fmt, ...);
void clear_error();
}  // namespace p4g

#define P4G_TRY(expr)                    \
    do {                                 \
        int p4g_e_ = (expr);             \
        if (p4g_e_) return p4g_e_;       \
    } while (0)

// --------------------------------------------------------------- options

namespace p4g {
// PETSc-style option list: "-key value" or "-key" (flag).  A token is a
// value unless it starts with '-' followed by a letter.
struct Options {
    std::vector<std::string> keys, vals;
    void parse(const char* s);
    bool has(const std::string& k) const;
    const std::string* get(const std::string& k) const;
    int         geti(const std::string& k, int def) const;
    double      getr(const std::string& k, double def) const;
    std::string gets(const std::string& k, const std::string& def) const;
    bool        getb(const std::string& k, bool def) const;
};
}  // namespace p4g

// ------------------------------------------------------------------- DMDA

struct p4g_da_s {
    int    dim = 2, mx = 3, my = 3, mz = 3;
    double L[3] = {1.0, 1.0, 1.0};
    double lo[3] = {0.0, 0.0, 0.0};   // bounding box: [lo, lo + L] per dimension
    int    refine = 0;
    DMDASNESFunctionFn* fn = nullptr;
    void*  fnctx = nullptr;
    DMDASNESJacobianFn* jac = nullptr;
    void*  jacctx = nullptr;
    void*  appctx = nullptr;
    std::vector<double> kcell;       // variable coefficient (finest cells)
    double flops = 0.0;              // PetscLogFlops total
    bool   internal = false;         // level grid owned by a solver
    long n() const { return (long)mx * (dim > 1 ? my : 1) * (dim > 2 ? mz : 1); }
    double h(int d) const {
        const int m = d == 0 ? mx : (d == 1 ? my : mz);
        return L[d] / (m - 1);
    }
};

// ------------------------------------------------- recording matrix (set-up)

// Rows recorded by MatSetValuesStencil are classified on the fly against the
// star-stencil pattern of Poisson{1,2,3}DJacobianLocal: boundary rows hold
// only a diagonal; interior rows hold a diagonal and couplings to interior
// neighbours only, the same value per direction.  The first row of each
// kind fixes the values; every later row must repeat them.
struct p4g_mat_s {
    p4g_da_s* da = nullptr;
    long   rows = 0, bad = 0;
    bool   have_b = false, have_i = false;
    double db = 0.0, d = 0.0;
    double s[3] = {0.0, 0.0, 0.0};   // couplings as magnitudes
    bool   have_s[3] = {false, false, false};
    std::string why;                 // first mismatch
    std::vector<char> seen;          // rows inserted
    bool   keep = false;             // also record every entry (MatIsSymmetric)
    std::vector<long> er, ec;        // recorded rows and columns, insertion order
    std::vector<double> ev;          // recorded values
};

// ------------------------------------------------------------ host vectors

struct p4g_vec_s {
    p4g_da_s* da = nullptr;
    std::vector<double> a;
    std::vector<double*> rows;       // DMDAVecGetArray views
    std::vector<double**> planes;
};

struct p4g_random_s {
    uint64_t seed = 0x12345678ull, counter = 0;
};

namespace p4g {
// pointer-of-pointer view over a host array of a DMDA (global indices)
void* make_view(p4g_da_s* da, double* base, std::vector<double*>& rows,
                std::vector<double**>& planes);
void  info_of(p4g_da_s* da, DMDALocalInfo* info);

// PETSc binary files: big-endian int32 / IEEE double, buffered writes
extern const int32_t kVecClassId, kMatClassId;
struct BinaryFile {
    FILE* fp = nullptr;
    std::vector<unsigned char> buf;
    bool failed = false;
    ~BinaryFile();
    bool open(const char* path, const char* mode);
    bool close();                       // flush and close; false on any write error
    bool flush();
    void bytes(const void* p, size_t k);
    void i32(int32_t v);
    void f64(double v);
    void f64s(const double* x, long n);
    bool read_i32(int32_t& v);
    bool read_f64s(double* x, long n);
    bool skip(long k);
};
}  // namespace p4g
