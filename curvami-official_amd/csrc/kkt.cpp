// Host KKT solver: [K J^T; J 0] by a sparse symmetric indefinite LDL^T (MKL
// PARDISO, mtype -2), with the symbolic analysis reused while the pattern stays
// the same and the inertia reported per factorization. The reference's solver
// is MumpsSymmetricSolver (core/include/curvami/solver/MumpsSymmetricSolver.h);
// the ABI is declared in include/curvami_kkt.h.
//
// The matrix goes to the factorization as the upper triangle in CSR: row i <
// n holds K(i, j) for j >= i (the transpose of the stored block-upper CSC,
// whose entries below the diagonal are dropped) and then J(r, i) at column
// n + r; row n + r holds its (zero) diagonal. A value map, built with the
// pattern, gathers K's stored values and sums J's triplets into that layout.
//
// The symmetric factorization pivots (1 x 1 and 2 x 2) within supernodes only,
// where MUMPS can also delay a pivot to an ancestor front; on a nearly singular
// saddle point (a flat folded sheet held by many constraint rows) that can
// leave a poor factor. Every factorization is therefore checked with one probe
// solve whose exact answer is known (b = A 1): a normwise backward error above
// 1e-12 makes the solver factor the full matrix by LU with weighted matching
// and use it for that factorization's solves, and the inertia is then reported
// as unknown (-1) instead of a count the poor factor cannot be trusted for.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/curvami_kkt.h"

// MKL PARDISO and service entry points (MKL reference manual; LP64 interface,
// 32-bit integers). The image ships MKL without its headers.
extern "C" {
void pardiso(void* pt, const int* maxfct, const int* mnum, const int* mtype, const int* phase, const int* n,
             const double* a, const int* ia, const int* ja, int* perm, const int* nrhs, int* iparm,
             const int* msglvl, double* b, double* x, int* error);
int MKL_Set_Num_Threads_Local(int nt);
int MKL_CBWR_Set(int settings);
}

// Conditional numerical reproducibility: the AVX2 code path in strict mode on
// every host with AVX2, and a fixed thread count for the factorization, make
// factorizations and solves bit-reproducible run to run and across hosts.
constexpr int kCbwrAvx2 = 10, kCbwrStrict = 0x10000;
constexpr int kCnrThreads = 8;
constexpr double kProbeBackwardError = 1e-12;

namespace {

const char* pardiso_message(int e) {
    switch (e) {
        case -1: return "input inconsistent";
        case -2: return "not enough memory";
        case -3: return "reordering problem";
        case -4: return "zero pivot, numerical factorization or iterative refinement problem";
        case -5: return "unclassified (internal) error";
        case -6: return "reordering failed";
        case -7: return "diagonal matrix is singular";
        case -8: return "32-bit integer overflow problem";
        case -9: return "not enough memory for OOC";
        case -10: return "error opening OOC files";
        case -11: return "read/write error with OOC files";
        default: return "error";
    }
}

// One solver handle: internal state, parameters, matrix type.
struct Handle {
    void* pt[64];
    int iparm[64];
    int mtype = -2;
    bool analysed = false;
    bool factored = false;
    explicit Handle(int mt) : mtype(mt) {
        std::memset(pt, 0, sizeof(pt));
        std::memset(iparm, 0, sizeof(iparm));
        iparm[0] = 1;    // no solver defaults
        iparm[1] = 2;    // nested-dissection fill-reducing ordering (METIS)
        iparm[7] = 2;    // at most two steps of iterative refinement
        iparm[9] = mt == -2 ? 8 : 13;   // pivot perturbation 1e-8 (symmetric), 1e-13 (LU)
        iparm[10] = 1;   // scaling
        iparm[12] = 1;   // weighted matching (saddle-point systems; without it the zero
                         // constraint block is factored with perturbed pivots)
        iparm[17] = -1;  // report the factor's entries
        iparm[20] = 1;   // symmetric: Bunch-Kaufman 1 x 1 and 2 x 2 pivots
        iparm[33] = kCnrThreads;   // reproducible parallel factorization with this many threads
        iparm[34] = 1;   // zero-based indices
    }
    int call(int nthreads, int phase, int N, const double* a, const int* ia, const int* ja, int nrhs, double* b,
             double* x) {
        const int maxfct = 1, mnum = 1, msglvl = 0;
        int error = 0;
        int idum = 0;
        if (nthreads > 0) MKL_Set_Num_Threads_Local(nthreads);
        pardiso(pt, &maxfct, &mnum, &mtype, &phase, &N, a, ia, ja, &idum, &nrhs, iparm, &msglvl, b, x, &error);
        if (nthreads > 0) MKL_Set_Num_Threads_Local(0);
        return error;
    }
    void release(int N, const int* ia, const int* ja) {
        if (analysed || factored) {
            double d = 0.0;
            call(0, -1, N, nullptr, ia, ja, 1, &d, &d);
        }
        analysed = factored = false;
    }
};

struct Kkt {
    Handle ldl{-2}, lu{11};
    int nthreads = 0;
    bool use_lu = false;        // the current factorization's solves go through the LU
    std::string err;
    // pattern as given (for the comparison that decides re-analysis)
    int n = 0, m = 0;
    std::vector<int> outer, inner, jr, jc;
    // upper-triangle CSR of the KKT matrix and the value map
    int N = 0;
    std::vector<int> ia, ja;
    std::vector<int64_t> ksrc;  // per CSR entry: index into K_values, or -1
    std::vector<int64_t> jpos;  // per J triplet: its CSR entry
    std::vector<double> a;
    // full CSR (both triangles) for the LU, built on first use: per entry its
    // upper-triangle entry
    std::vector<int> ia2, ja2;
    std::vector<int64_t> src2;
    std::vector<double> a2;
    int64_t analyses = 0, factorizations = 0, solves = 0, lu_fallbacks = 0;
    int64_t factor_nnz = 0, perturbed = 0, t_analysis_us = 0, t_factor_us = 0;
    double probe_berr = 0.0;

    void release() {
        if (!ia.empty()) ldl.release(N, ia.data(), ja.data());
        if (!ia2.empty()) lu.release(N, ia2.data(), ja2.data());
        use_lu = false;
    }
    ~Kkt() { release(); }

    // y = A x for the symmetric matrix held as its upper triangle
    void symv(const double* x, double* y) const {
        std::fill(y, y + N, 0.0);
        for (int i = 0; i < N; ++i)
            for (int q = ia[i]; q < ia[i + 1]; ++q) {
                const int j = ja[q];
                y[i] += a[q] * x[j];
                if (j != i) y[j] += a[q] * x[i];
            }
    }
    double norm_inf() const {
        std::vector<double> r(N, 0.0);
        for (int i = 0; i < N; ++i)
            for (int q = ia[i]; q < ia[i + 1]; ++q) {
                r[i] += std::fabs(a[q]);
                if (ja[q] != i) r[ja[q]] += std::fabs(a[q]);
            }
        return N ? *std::max_element(r.begin(), r.end()) : 0.0;
    }
    void build_full() {
        std::vector<int> cnt(N + 1, 0);
        for (int i = 0; i < N; ++i)
            for (int q = ia[i]; q < ia[i + 1]; ++q) { cnt[i + 1]++; if (ja[q] != i) cnt[ja[q] + 1]++; }
        ia2.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) ia2[i + 1] = ia2[i] + cnt[i + 1];
        ja2.assign(ia2[N], 0);
        src2.assign(ia2[N], 0);
        std::vector<int> fill(ia2.begin(), ia2.end() - 1);
        // row r of the full matrix: the mirrored entries (columns < r) come from
        // rows i < r in ascending i, then row r's own upper entries: ascending
        for (int i = 0; i < N; ++i)
            for (int q = ia[i]; q < ia[i + 1]; ++q) {
                const int j = ja[q];
                if (j == i) continue;
                const int p = fill[j]++;
                ja2[p] = i; src2[p] = q;
            }
        for (int i = 0; i < N; ++i)
            for (int q = ia[i]; q < ia[i + 1]; ++q) {
                const int p = fill[i]++;
                ja2[p] = ja[q]; src2[p] = q;
            }
        a2.assign(ia2[N], 0.0);
    }
};

int64_t us_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct cvm_kkt {
    Kkt k;
};

extern "C" {

int cvm_kkt_create(int nthreads, cvm_kkt** out) {
    if (!out) return 1;
    // MKL dlopens its CPU-specific kernels, which resolve symbols of the core
    // and threading libraries in the global scope: a host that loaded this
    // library with local scope (a ctypes CDLL) would leave them unresolved, so
    // the libraries already loaded as its dependencies are made global
    for (const char* lib : {"libmkl_core.so.1", "libmkl_gnu_thread.so.1", "libmkl_intel_lp64.so.1"})
        (void)dlopen(lib, RTLD_NOW | RTLD_GLOBAL | RTLD_NOLOAD);
    static const int cbwr = MKL_CBWR_Set(kCbwrAvx2 | kCbwrStrict);   // before the first MKL computation
    (void)cbwr;
    *out = new (std::nothrow) cvm_kkt();
    if (!*out) return 6;
    (*out)->k.nthreads = nthreads > 0 ? nthreads : 0;
    return 0;
}

void cvm_kkt_destroy(cvm_kkt* s) { delete s; }

const char* cvm_kkt_last_error(const cvm_kkt* s) { return s ? s->k.err.c_str() : "null solver"; }

int cvm_kkt_set_pattern(cvm_kkt* s, int n, const int* outer, const int* inner, int m, int64_t nt, const int* rows,
                        const int* cols) {
    if (!s) return 1;
    Kkt& k = s->k;
    if (n <= 0 || m < 0 || nt < 0 || !outer || outer[0] != 0 || (outer[n] > 0 && !inner) ||
        (nt > 0 && (!rows || !cols))) {
        k.err = "bad pattern arrays";
        return 1;
    }
    const int64_t nnzk = outer[n];
    const bool same = n == k.n && m == k.m && static_cast<int64_t>(k.inner.size()) == nnzk &&
                      static_cast<int64_t>(k.jr.size()) == nt && !k.ia.empty() &&
                      std::equal(outer, outer + n + 1, k.outer.begin()) &&
                      std::equal(inner, inner + nnzk, k.inner.begin()) && std::equal(rows, rows + nt, k.jr.begin()) &&
                      std::equal(cols, cols + nt, k.jc.begin());
    if (same) return 0;
    for (int64_t t = 0; t < nt; ++t)
        if (rows[t] < 0 || rows[t] >= m || cols[t] < 0 || cols[t] >= n) {
            k.err = "constraint triplet outside the system";
            return 1;
        }
    for (int j = 0; j < n; ++j)
        if (outer[j + 1] < outer[j]) {
            k.err = "K column pointers decrease";
            return 1;
        }
    if (static_cast<int64_t>(n) + m > 0x7fffffffLL || nnzk + nt + n + m > 0x7fffffffLL) {
        k.err = "KKT system exceeds 32-bit indices";
        return 1;
    }
    k.release();
    k.ia.clear();           // no valid layout until this pattern is complete
    k.ia2.clear(); k.ja2.clear(); k.src2.clear(); k.a2.clear();
    k.n = n; k.m = m;
    k.outer.assign(outer, outer + n + 1);
    k.inner.assign(inner, inner + nnzk);
    k.jr.assign(rows, rows + nt);
    k.jc.assign(cols, cols + nt);
    k.N = n + m;
    // per-row counts: upper K entries (row i <= column j), a diagonal when K
    // stores none, distinct J columns n + r per DOF row, the constraint diagonals
    std::vector<int> cnt(k.N + 1, 0);
    std::vector<char> hasdiag(n, 0);
    for (int j = 0; j < n; ++j)
        for (int64_t p = outer[j]; p < outer[j + 1]; ++p) {
            const int i = inner[p];
            if (i < 0 || i >= n) {
                k.err = "K row index outside the system";
                return 1;
            }
            if (i <= j) { cnt[i + 1]++; if (i == j) hasdiag[i] = 1; }
        }
    for (int i = 0; i < n; ++i) if (!hasdiag[i]) cnt[i + 1]++;
    // J entries per DOF row, duplicates merged: sort triplet indices by (col, row)
    std::vector<int64_t> ord(nt);
    for (int64_t t = 0; t < nt; ++t) ord[t] = t;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        return cols[a] != cols[b] ? cols[a] < cols[b] : rows[a] < rows[b];
    });
    for (int64_t q = 0; q < nt; ++q) {
        const int64_t t = ord[q];
        if (q == 0 || cols[ord[q - 1]] != cols[t] || rows[ord[q - 1]] != rows[t]) cnt[cols[t] + 1]++;
    }
    for (int r = 0; r < m; ++r) cnt[n + r + 1] = 1;
    k.ia.assign(k.N + 1, 0);
    for (int i = 0; i < k.N; ++i) k.ia[i + 1] = k.ia[i] + cnt[i + 1];
    const int64_t tot = k.ia[k.N];
    k.ja.assign(tot, 0);
    k.ksrc.assign(tot, -1);
    k.jpos.assign(nt, -1);
    std::vector<int> fill(k.ia.begin(), k.ia.end() - 1);
    // K part: columns j ascending give each row's columns in ascending order; a
    // missing diagonal goes first (every other entry of row i has column > i)
    for (int i = 0; i < n; ++i) if (!hasdiag[i]) { k.ja[fill[i]] = i; fill[i]++; }
    for (int j = 0; j < n; ++j)
        for (int64_t p = outer[j]; p < outer[j + 1]; ++p) {
            const int i = inner[p];
            if (i > j) continue;
            const int q = fill[i]++;
            k.ja[q] = j;
            k.ksrc[q] = p;
        }
    for (int64_t q = 0; q < nt; ++q) {
        const int64_t t = ord[q];
        const bool fresh = q == 0 || cols[ord[q - 1]] != cols[t] || rows[ord[q - 1]] != rows[t];
        if (fresh) {
            const int pos = fill[cols[t]]++;
            k.ja[pos] = n + rows[t];
        }
        k.jpos[t] = fill[cols[t]] - 1;
    }
    for (int r = 0; r < m; ++r) k.ja[fill[n + r]++] = n + r;
    // a column appearing twice in a row (a duplicate K entry) is refused
    for (int i = 0; i < k.N; ++i)
        for (int q = k.ia[i] + 1; q < k.ia[i + 1]; ++q)
            if (k.ja[q] <= k.ja[q - 1]) {
                k.err = "K pattern holds a duplicate or unsorted entry";
                k.ia.clear();
                return 1;
            }
    k.a.assign(tot, 0.0);
    return 0;
}

int cvm_kkt_factorize(cvm_kkt* s, const double* K_values, const double* J_values, int* neg, int* pos) {
    if (!s) return 1;
    Kkt& k = s->k;
    if (k.ia.empty()) { k.err = "no pattern (cvm_kkt_set_pattern)"; return 7; }
    if (!K_values || (!k.jr.empty() && !J_values)) { k.err = "null values"; return 1; }
    const int64_t tot = k.ia[k.N];
    for (int64_t q = 0; q < tot; ++q) k.a[q] = k.ksrc[q] >= 0 ? K_values[k.ksrc[q]] : 0.0;
    for (size_t t = 0; t < k.jpos.size(); ++t) k.a[k.jpos[t]] += J_values[t];
    k.use_lu = false;
    auto fail = [&](const char* what, int e) {
        k.err = std::string("KKT ") + what + " failed (PARDISO error " + std::to_string(e) + ": " +
                pardiso_message(e) + ")";
        return CVM_E_KKT;
    };
    double d = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    if (!k.ldl.analysed) {
        const int e = k.ldl.call(k.nthreads, 11, k.N, k.a.data(), k.ia.data(), k.ja.data(), 1, &d, &d);
        if (e != 0) { k.ldl.release(k.N, k.ia.data(), k.ja.data()); return fail("analysis", e); }
        k.ldl.analysed = true;
        ++k.analyses;
        k.t_analysis_us = us_since(t0);
        t0 = std::chrono::steady_clock::now();
    }
    int e = k.ldl.call(k.nthreads, 22, k.N, k.a.data(), k.ia.data(), k.ja.data(), 1, &d, &d);
    k.ldl.factored = e == 0;
    if (e != 0) return fail("factorization", e);
    ++k.factorizations;
    k.factor_nnz = k.ldl.iparm[17];
    k.perturbed = k.ldl.iparm[13];
    int npos = k.ldl.iparm[21], nneg = k.ldl.iparm[22];
    // probe: b = A 1 must give x = 1
    std::vector<double> one(k.N, 1.0), b(k.N), x(k.N), r(k.N);
    k.symv(one.data(), b.data());
    e = k.ldl.call(k.nthreads, 33, k.N, k.a.data(), k.ia.data(), k.ja.data(), 1, b.data(), x.data());
    double berr = 1.0;
    if (e == 0) {
        k.symv(x.data(), r.data());
        double rn = 0.0, xn = 0.0, bn = 0.0;
        for (int i = 0; i < k.N; ++i) {
            rn = std::max(rn, std::fabs(r[i] - b[i]));
            xn = std::max(xn, std::fabs(x[i]));
            bn = std::max(bn, std::fabs(b[i]));
        }
        const double den = k.norm_inf() * xn + bn;
        berr = den > 0.0 ? rn / den : 0.0;
        if (!std::isfinite(berr)) berr = 1.0;
    }
    k.probe_berr = berr;
    if (berr > kProbeBackwardError) {
        // the LU of the full matrix for this factorization's solves
        if (k.ia2.empty()) k.build_full();
        for (size_t q = 0; q < k.a2.size(); ++q) k.a2[q] = k.a[k.src2[q]];
        if (!k.lu.analysed) {
            e = k.lu.call(k.nthreads, 11, k.N, k.a2.data(), k.ia2.data(), k.ja2.data(), 1, &d, &d);
            if (e != 0) { k.lu.release(k.N, k.ia2.data(), k.ja2.data()); return fail("LU analysis", e); }
            k.lu.analysed = true;
        }
        e = k.lu.call(k.nthreads, 22, k.N, k.a2.data(), k.ia2.data(), k.ja2.data(), 1, &d, &d);
        k.lu.factored = e == 0;
        if (e != 0) return fail("LU factorization", e);
        k.use_lu = true;
        ++k.lu_fallbacks;
        npos = nneg = -1;
    }
    k.t_factor_us = us_since(t0);
    if (pos) *pos = npos;
    if (neg) *neg = nneg;
    return 0;
}

int cvm_kkt_solve(cvm_kkt* s, int nrhs, const double* rhs, double* x) {
    if (!s) return 1;
    Kkt& k = s->k;
    if (!k.ldl.factored || (k.use_lu && !k.lu.factored)) { k.err = "no factorization (cvm_kkt_factorize)"; return 7; }
    if (nrhs <= 0 || !rhs || !x) { k.err = "bad right-hand sides"; return 1; }
    std::vector<double> b(rhs, rhs + static_cast<size_t>(k.N) * nrhs);
    const int e = k.use_lu ? k.lu.call(k.nthreads, 33, k.N, k.a2.data(), k.ia2.data(), k.ja2.data(), nrhs, b.data(), x)
                           : k.ldl.call(k.nthreads, 33, k.N, k.a.data(), k.ia.data(), k.ja.data(), nrhs, b.data(), x);
    if (e != 0) {
        k.err = std::string("KKT solve failed (PARDISO error ") + std::to_string(e) + ": " + pardiso_message(e) + ")";
        return CVM_E_KKT;
    }
    ++k.solves;
    return 0;
}

int cvm_kkt_stats(const cvm_kkt* s, int64_t* out10) {
    if (!s || !out10) return 1;
    const Kkt& k = s->k;
    out10[0] = k.analyses; out10[1] = k.factorizations; out10[2] = k.solves;
    out10[3] = k.ia.empty() ? 0 : k.ia[k.N];
    out10[4] = k.factor_nnz; out10[5] = k.perturbed;
    out10[6] = k.t_analysis_us; out10[7] = k.t_factor_us;
    out10[8] = k.lu_fallbacks;
    out10[9] = static_cast<int64_t>(std::llround(-std::log10(std::max(k.probe_berr, 1e-300))));
    return 0;
}

}  // extern "C"
