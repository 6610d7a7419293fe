// Dense solves of the small problems of a batch of independent bubbles, all
// in one launch: one work-group per problem factors its FP64 G block in place
// with partial pivoting and solves G x = b (Simulation::solve_system with
// PartialPivLU, Simulation.cpp:40-71, for a batch of problems at once).
//
// The blocked right-looking elimination keeps an 8-column panel and the
// matching 8 rows of U in LDS: the panel is factored there (pivot search over
// the column, interchanges of whole rows as LAPACK getrf / PartialPivLU do,
// ties to the smallest row), the 8 U rows of the trailing columns are solved
// with the panel's unit-lower block, and the trailing matrix takes one rank-8
// update (one read and one write of each entry per panel).  The triangular
// solves follow in the same work-group.  Many problems of a few hundred
// unknowns are latency bound when solved one after another with the blocked
// multi-launch LU; here every problem is a work-group of one launch.

#include "internal.hpp"

namespace bemhip {

namespace {

constexpr int ST = 1024;       // threads per work-group
constexpr int SW = 8;          // panel width
constexpr int SMAX = SMALL_MAX_N;

static_assert(SMAX <= ST, "one row per thread in the panel kernels");

// first maximum of |a| (smallest row on ties) over the work-group; v < 0 means
// "no candidate"
__device__ inline void better(double& v, int& i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__global__ __launch_bounds__(ST) void k_small_lu_solve(const SmallJob* __restrict__ jobs, double* G,
                                                       const double* __restrict__ b_all, double* __restrict__ x_all,
                                                       int* __restrict__ info) {
    // factorisation: panel (SW columns x SMAX rows) and U rows (SW x SMAX);
    // solves: right-hand side and solution (the same storage)
    __shared__ double buf[2 * SW * SMAX];
    __shared__ int piv[SMAX];
    __shared__ double red_v[ST / 64];
    __shared__ int red_i[ST / 64];
    __shared__ int s_piv, s_info;
    double* Lp = buf;              // Lp[c * SMAX + i]: panel column c, row i
    double* Up = buf + SW * SMAX;  // Up[c * SMAX + j]: U row k0 + c, column j

    const SmallJob jb = jobs[blockIdx.x];
    const int n = jb.n;
    double* A = G + jb.gofs;       // column-major n x n
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) s_info = 0;

    for (int k0 = 0; k0 < n; k0 += SW) {
        const int w = min(SW, n - k0);
        for (int e = tid; e < w * SMAX; e += ST) {
            const int c = e / SMAX, i = e % SMAX;
            if (i >= k0 && i < n) Lp[c * SMAX + i] = A[(size_t)(k0 + c) * n + i];
        }
        __syncthreads();
        for (int c = 0; c < w; ++c) {
            const int col = k0 + c;
            // pivot: first maximum of |a| in rows col .. n-1 (one row per thread)
            double v = -1.0;
            int vi = n;
            if (tid >= col && tid < n) { v = fabs(Lp[c * SMAX + tid]); vi = tid; }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double v2 = __shfl_xor(v, off);
                const int i2 = __shfl_xor(vi, off);
                better(v, vi, v2, i2);
            }
            if (lane == 0) { red_v[wv] = v; red_i[wv] = vi; }
            __syncthreads();
            if (tid == 0) {
                double bv = red_v[0];
                int bi = red_i[0];
                for (int q = 1; q < ST / 64; ++q) better(bv, bi, red_v[q], red_i[q]);
                s_piv = bi < n ? bi : col;
                piv[col] = s_piv;
            }
            __syncthreads();
            const int p = s_piv;
            if (p != col) {
                // interchange rows col and p over the whole matrix: the panel
                // in LDS, every other column in place
                for (int j = tid; j < n; j += ST) {
                    if (j >= k0 && j < k0 + w) {
                        const int cc = j - k0;
                        const double t = Lp[cc * SMAX + col];
                        Lp[cc * SMAX + col] = Lp[cc * SMAX + p];
                        Lp[cc * SMAX + p] = t;
                    } else {
                        double* a = A + (size_t)j * n;
                        const double t = a[col];
                        a[col] = a[p];
                        a[p] = t;
                    }
                }
                __syncthreads();
            }
            const double d = Lp[c * SMAX + col];
            if (d != 0.0) {
                // multipliers and the rank-1 update of the panel's later columns
                const int i = tid;
                if (i > col && i < n) {
                    const double l = Lp[c * SMAX + i] / d;
                    Lp[c * SMAX + i] = l;
                    for (int cc = c + 1; cc < w; ++cc) Lp[cc * SMAX + i] -= l * Lp[cc * SMAX + col];
                }
            } else if (tid == 0 && s_info == 0) {
                s_info = col + 1;   // zero pivot (LAPACK info): no multipliers, elimination goes on
            }
            __syncthreads();
        }
        // the factored panel back to its columns
        for (int e = tid; e < w * SMAX; e += ST) {
            const int c = e / SMAX, i = e % SMAX;
            if (i >= k0 && i < n) A[(size_t)(k0 + c) * n + i] = Lp[c * SMAX + i];
        }
        const int j0 = k0 + w;
        if (j0 < n) {
            // U rows k0 .. k0+w-1 of the trailing columns: L11^-1 A12 (one column per thread)
            const int j = j0 + tid;
            if (j < n) {
                double u[SW];
                double* a = A + (size_t)j * n + k0;
#pragma unroll
                for (int r = 0; r < SW; ++r) u[r] = r < w ? a[r] : 0.0;
#pragma unroll
                for (int r = 1; r < SW; ++r)
#pragma unroll
                    for (int s = 0; s < r; ++s)
                        if (r < w) u[r] -= Lp[s * SMAX + k0 + r] * u[s];
#pragma unroll
                for (int r = 0; r < SW; ++r)
                    if (r < w) { a[r] = u[r]; Up[r * SMAX + j] = u[r]; }
            }
            __syncthreads();
            // rank-w update of the trailing block: 256 row lanes x 4 column groups,
            // each thread keeps its row's multipliers in registers
            const int ri = tid & 255, cg = tid >> 8;
            for (int i = j0 + ri; i < n; i += 256) {
                double l[SW];
#pragma unroll
                for (int c = 0; c < SW; ++c) l[c] = c < w ? Lp[c * SMAX + i] : 0.0;
                for (int jj = j0 + cg; jj < n; jj += 4) {
                    double* a = A + (size_t)jj * n + i;
                    double s = *a;
#pragma unroll
                    for (int c = 0; c < SW; ++c) s -= l[c] * Up[c * SMAX + jj];
                    *a = s;
                }
            }
        }
        __syncthreads();
    }

    // solve: x = U^-1 L^-1 P b
    double* xs = buf;          // right-hand side, then L^-1 P b
    double* xo = buf + SMAX;   // solution
    const double* b = b_all + jb.v0;
    for (int i = tid; i < n; i += ST) xs[i] = b[i];
    __syncthreads();
    if (tid == 0) {
        for (int c = 0; c < n; ++c) {
            const int p = piv[c];
            if (p != c) { const double t = xs[c]; xs[c] = xs[p]; xs[p] = t; }
        }
    }
    __syncthreads();
    for (int c = 0; c < n - 1; ++c) {
        const double y = xs[c];
        const double* lc = A + (size_t)c * n;
        for (int i = c + 1 + tid; i < n; i += ST) xs[i] -= lc[i] * y;
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        const double* uc = A + (size_t)c * n;
        const double xc = xs[c] / uc[c];
        if (tid == 0) xo[c] = xc;
        for (int i = tid; i < c; i += ST) xs[i] -= uc[i] * xc;
        __syncthreads();
    }
    double* x = x_all + jb.v0;
    for (int i = tid; i < n; i += ST) x[i] = xo[i];
    if (tid == 0) info[blockIdx.x] = s_info;
}

}  // namespace

void small_solve_launch(hipStream_t st, const SmallJob* d_jobs, int count, double* G64, const double* b, double* x,
                        int* d_info) {
    if (count <= 0) return;
    k_small_lu_solve<<<count, ST, 0, st>>>(d_jobs, G64, b, x, d_info);
    BH_HIP(hipGetLastError());
}

}  // namespace bemhip
