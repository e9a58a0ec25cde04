/*
 * curvami_kkt.h -- C ABI of the host KKT solver: the symmetric indefinite
 * saddle-point system [K J^T; J 0] of Curvami's Newton and arc-length loops,
 * solved by a sparse LDL^T (MKL PARDISO, symmetric indefinite) whose symbolic
 * analysis is kept while the pattern stays the same and whose negative-pivot
 * count (the KKT inertia) is reported with every factorization.
 *
 * Replaces MumpsSymmetricSolver (core/include/curvami/solver/MumpsSymmetricSolver.h):
 * sym = 2 factorization of the block system (:46-55), symbolic analysis (job 1)
 * only when the pattern changes (:184-196), numerical factorization (job 2,
 * :198-220) and solve (job 3, :221-230) per call, INFOG(12) as the negative
 * pivot count (:58-60, :217). Host memory only; no GPU work. Factorizations
 * and solves are bit-reproducible run to run, across thread counts and across
 * hosts with AVX2 (the solver's conditional-reproducibility mode).
 *
 * Conventions as curvami_gpu.h: plain C types, int status codes (CVM_OK = 0,
 * CVM_E_INVALID = 1, CVM_E_STATE = 7, CVM_E_KKT = 9: the factorization or
 * solve failed, the reference's last_error_ != 0), never exceptions. A solver
 * object is not thread-safe; many may coexist.
 */
#ifndef CURVAMI_KKT_H
#define CURVAMI_KKT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVM_KKT_API __attribute__((visibility("default")))
#define CVM_E_KKT 9

typedef struct cvm_kkt cvm_kkt;

/* nthreads <= 0: the runtime's default thread count. */
CVM_KKT_API int cvm_kkt_create(int nthreads, cvm_kkt** out);
CVM_KKT_API void cvm_kkt_destroy(cvm_kkt* s);
CVM_KKT_API const char* cvm_kkt_last_error(const cvm_kkt* s);

/* Pattern of the system: K (n x n) in the assembly's block-upper CSC
 * (outer[n + 1], inner[outer[n]]; entries below the diagonal, which the stored
 * 3 x 3 diagonal node blocks hold, are ignored, as MUMPS sym = 2 ignores them,
 * MumpsSymmetricSolver.h:84-126) and J (m x n) as triplets rows[nt], cols[nt]
 * (duplicates are summed, as Eigen's setFromTriplets does). The next
 * factorization re-runs the symbolic analysis only if this pattern differs from
 * the analysed one (MumpsSymmetricSolver.h:184-196). */
CVM_KKT_API int cvm_kkt_set_pattern(cvm_kkt* s, int n, const int* outer, const int* inner, int m, int64_t nt,
                                    const int* rows, const int* cols);

/* Numerical factorization with K_values[outer[n]] (the pattern's order) and
 * J_values[nt]. neg / pos (may be NULL): negative and positive pivots of the
 * KKT matrix (INFOG(12) and its complement). Each factorization is checked by
 * a probe solve with a known answer; when its normwise backward error exceeds
 * 1e-12 (a nearly singular saddle point that pivoting within supernodes cannot
 * factor stably), the full matrix is also factored by LU with weighted
 * matching, the solves use it, and neg = pos = -1 (inertia unknown).
 * CVM_E_KKT on failure (the message names the solver's error code). */
CVM_KKT_API int cvm_kkt_factorize(cvm_kkt* s, const double* K_values, const double* J_values, int* neg, int* pos);

/* Solve with the last factorization: nrhs right-hand sides of n + m entries,
 * column-major, into x (x may equal rhs). */
CVM_KKT_API int cvm_kkt_solve(cvm_kkt* s, int nrhs, const double* rhs, double* x);

/* out10: analyses, factorizations, solves, stored entries of the KKT upper
 * triangle, entries of the factor, perturbed pivots of the last factorization,
 * last analysis and last factorization time (microseconds), factorizations
 * that fell back to the LU, -log10 of the last probe's backward error. */
CVM_KKT_API int cvm_kkt_stats(const cvm_kkt* s, int64_t* out10);

#ifdef __cplusplus
}
#endif
#endif /* CURVAMI_KKT_H */
