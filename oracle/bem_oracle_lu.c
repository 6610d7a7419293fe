/*
 * bem_oracle_lu.c -- cache-blocked GEPP for the CPU baseline (TEST
 * INFRASTRUCTURE ONLY: the benchmark's CPU leg times it; the product path
 * never links it).
 *
 * The reference factors G with Eigen's PartialPivLU (Simulation.cpp:57-61),
 * a blocked right-looking LU whose trailing update is Eigen's packed,
 * register-blocked GEMM.  This restates that structure: panels of NB columns
 * factored by unblocked GEPP (first maximum wins, as bo_lu_factor), the row
 * interchanges applied across the matrix, U12 = L11^-1 A12 by column-parallel
 * forward substitution and A22 -= L21 U12 as a packed GEMM: A in 8-row
 * slivers and B in 4-column slivers, an 8 x 4 register tile of vector
 * accumulators per micro-kernel call, OpenMP over tiles of C.
 *
 * max_panels > 0 stops after that many panels (a timed sample of a large
 * factorisation); the return value is the LAPACK info of what was factored.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double v4d __attribute__((vector_size(32)));

#define MC 128   /* rows of A packed per block (16 slivers of 8)   */
#define NC 256   /* columns of B packed per block (64 slivers of 4) */

/* C (8 x 4, ld ldc) -= Ap (k x 8 sliver) Bp (k x 4 sliver) */
static inline void micro_8x4(int64_t k, const double* restrict Ap, const double* restrict Bp, double* restrict C,
                             int64_t ldc, int64_t mr, int64_t nr) {
    v4d c00 = {0}, c01 = {0}, c02 = {0}, c03 = {0}, c10 = {0}, c11 = {0}, c12 = {0}, c13 = {0};
    for (int64_t p = 0; p < k; ++p) {
        v4d a0, a1;
        memcpy(&a0, Ap + 8 * p, 32);
        memcpy(&a1, Ap + 8 * p + 4, 32);
        const double* b = Bp + 4 * p;
        const v4d b0 = {b[0], b[0], b[0], b[0]}, b1 = {b[1], b[1], b[1], b[1]};
        const v4d b2 = {b[2], b[2], b[2], b[2]}, b3 = {b[3], b[3], b[3], b[3]};
        c00 += a0 * b0; c10 += a1 * b0;
        c01 += a0 * b1; c11 += a1 * b1;
        c02 += a0 * b2; c12 += a1 * b2;
        c03 += a0 * b3; c13 += a1 * b3;
    }
    const v4d cc[4][2] = {{c00, c10}, {c01, c11}, {c02, c12}, {c03, c13}};
    for (int64_t j = 0; j < nr; ++j) {
        double t[8];
        memcpy(t, &cc[j][0], 32);
        memcpy(t + 4, &cc[j][1], 32);
        double* cj = C + (size_t)j * ldc;
        for (int64_t i = 0; i < mr; ++i) cj[i] -= t[i];
    }
}

/* C (m x n, ld ldc) -= A (m x k, ld lda) B (k x n, ld ldb), column-major */
static void gemm_minus(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                       double* C, int64_t ldc, int nthreads) {
    const int64_t nbi = (m + MC - 1) / MC, nbj = (n + NC - 1) / NC;
#pragma omp parallel num_threads(nthreads)
    {
        double* Ap = (double*)aligned_alloc(64, sizeof(double) * MC * k + 64);
        double* Bp = (double*)aligned_alloc(64, sizeof(double) * NC * k + 64);
#pragma omp for collapse(2) schedule(dynamic, 1)
        for (int64_t bj = 0; bj < nbj; ++bj)
            for (int64_t bi = 0; bi < nbi; ++bi) {
                const int64_t j0 = bj * NC, nc = (n - j0 < NC) ? n - j0 : NC;
                const int64_t i0 = bi * MC, mc = (m - i0 < MC) ? m - i0 : MC;
                /* pack B block: 4-column slivers, k-major */
                for (int64_t s = 0; s < (nc + 3) / 4; ++s)
                    for (int64_t p = 0; p < k; ++p)
                        for (int64_t j = 0; j < 4; ++j) {
                            const int64_t jj = 4 * s + j;
                            Bp[(size_t)s * 4 * k + 4 * p + j] = jj < nc ? B[(size_t)(j0 + jj) * ldb + p] : 0.0;
                        }
                /* pack A block: 8-row slivers, k-major */
                for (int64_t s = 0; s < (mc + 7) / 8; ++s)
                    for (int64_t p = 0; p < k; ++p)
                        for (int64_t i = 0; i < 8; ++i) {
                            const int64_t ii = 8 * s + i;
                            Ap[(size_t)s * 8 * k + 8 * p + i] = ii < mc ? A[(size_t)p * lda + i0 + ii] : 0.0;
                        }
                for (int64_t sj = 0; sj < (nc + 3) / 4; ++sj)
                    for (int64_t si = 0; si < (mc + 7) / 8; ++si) {
                        const int64_t mr = (mc - 8 * si < 8) ? mc - 8 * si : 8;
                        const int64_t nr = (nc - 4 * sj < 4) ? nc - 4 * sj : 4;
                        micro_8x4(k, Ap + (size_t)si * 8 * k, Bp + (size_t)sj * 4 * k,
                                  C + (size_t)(j0 + 4 * sj) * ldc + i0 + 8 * si, ldc, mr, nr);
                    }
            }
        free(Ap);
        free(Bp);
    }
}

int64_t bo_lu_factor_blocked(int64_t N, double* A, int64_t* piv, int nthreads, int64_t NB, int64_t max_panels) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    if (NB <= 0) NB = 256;
    int64_t info = 0, panels = 0;
#define A_(i, j) A[(size_t)(j) * N + (i)]
    for (int64_t k0 = 0; k0 < N; k0 += NB) {
        if (max_panels > 0 && panels++ >= max_panels) break;
        const int64_t kb = (N - k0 < NB) ? N - k0 : NB;
        const int64_t k1 = k0 + kb;
        /* panel: unblocked GEPP on columns k0..k1 (rows k0..N), first maximum wins */
        for (int64_t k = k0; k < k1; ++k) {
            int64_t p = k;
            double best = fabs(A_(k, k));
            for (int64_t i = k + 1; i < N; ++i) {
                const double v = fabs(A_(i, k));
                if (v > best) { best = v; p = i; }
            }
            piv[k] = p;
            if (p != k)
                for (int64_t j = k0; j < k1; ++j) { const double t = A_(k, j); A_(k, j) = A_(p, j); A_(p, j) = t; }
            const double d = A_(k, k);
            if (d == 0.0) { if (!info) info = k + 1; continue; }
            const double rd = 1.0 / d;
#pragma omp parallel for num_threads(nthreads) schedule(static) if (N - k > 4096)
            for (int64_t i = k + 1; i < N; ++i) A_(i, k) *= rd;
            for (int64_t j = k + 1; j < k1; ++j) {
                const double u = A_(k, j);
                if (u == 0.0) continue;
#pragma omp parallel for num_threads(nthreads) schedule(static) if (N - k > 4096)
                for (int64_t i = k + 1; i < N; ++i) A_(i, j) -= A_(i, k) * u;
            }
        }
        /* the panel's interchanges across the other columns */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = 0; j < N; ++j) {
            if (j >= k0 && j < k1) continue;
            for (int64_t k = k0; k < k1; ++k) {
                const int64_t p = piv[k];
                if (p != k) { const double t = A_(k, j); A_(k, j) = A_(p, j); A_(p, j) = t; }
            }
        }
        if (k1 >= N) break;
        /* U12 = L11^-1 A12, a column per task */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = k1; j < N; ++j)
            for (int64_t k = k0; k < k1; ++k) {
                const double u = A_(k, j);
                if (u == 0.0) continue;
                for (int64_t i = k + 1; i < k1; ++i) A_(i, j) -= A_(i, k) * u;
            }
        /* A22 -= L21 U12 */
        gemm_minus(N - k1, N - k1, kb, &A_(k1, k0), N, &A_(k0, k1), N, &A_(k1, k1), N, nthreads);
    }
#undef A_
    return info;
}
