// Dense solvers for G psi = H phi on gfx950.
//
// Reference: Simulation::solve_system (Simulation.cpp:40-71) with
// Eigen::PartialPivLU (LU with partial pivoting, first maximum wins) or
// Eigen::BiCGSTAB (diagonal preconditioner, tol = machine epsilon, maxit = 2N).
//
// LU design (column-major, in place, pivots as absolute row numbers):
//   * blocked right-looking, outer block NB = 256: panel, row interchanges on
//     both sides, TRSM of the block row, trailing GEMM update;
//   * the 256-wide panel is factored recursively (halves down to 32 columns);
//   * a 32-column leaf is one cooperative kernel: every work-group keeps up to
//     256 rows of the leaf in LDS, and each column costs one grid barrier
//     (candidate pivot rows and the current diagonal row are published before
//     it, so the interchange needs no second barrier);
//   * trailing updates are plain GEMMs (rocBLAS: FP64 / FP32 MFMA), or, for the
//     split-fp16 variant, one fp16 MFMA GEMM with K' = 3K on [Ah Ah Al] x
//     [Bh; Bl; Bh] accumulating in FP32 (block-scaled by powers of two).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace bemhip {

constexpr int LEAF = 32;     // cooperative leaf width
constexpr int PROWS = 256;   // rows per panel work-group (one per thread)
constexpr int MAXP = 256;    // max panel work-groups -> N <= 65536

// panel workspace layout (bytes)
struct PanelWs {
    static constexpr size_t BAR = 0;           // uint32 barrier counter
    static constexpr size_t ERR = 4;           // uint32 flags: 1 barrier timeout, 2 zero pivot
    static constexpr size_t ZCOL = 8;          // int32 first zero-pivot column (+1)
    static constexpr size_t CVAL = 64;                                  // double [2][MAXP]
    static constexpr size_t CIDX = CVAL + 2 * MAXP * sizeof(double);    // int [2][MAXP]
    static constexpr size_t CROW = CIDX + 2 * MAXP * sizeof(int);       // double [2][MAXP][LEAF]
    static constexpr size_t DROW = CROW + 2 * MAXP * LEAF * sizeof(double);  // double [2][LEAF]
    static constexpr size_t BYTES = DROW + 2 * LEAF * sizeof(double);
};

__device__ inline void st_relaxed(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_relaxed(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_relaxed(const int* p) {
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// all work-groups of a cooperative launch; bounded spin sets an error flag
__device__ inline void grid_barrier(unsigned int* bar, unsigned int target, unsigned int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __atomic_thread_fence(__ATOMIC_RELEASE);  // agent scope via HIP default
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned int spins = 0;
        // after any timeout the flag stays set and no later barrier waits, so
        // a failed launch drains quickly and the host reports the error
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target &&
               !(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u)) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 25)) {
                __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
    __syncthreads();
}

// (value, index, tag) argmax over a 256-thread block; ties -> smallest index
__device__ inline void block_argmax(double& v, int& i, int& tag, double* sv, int* si, int* st) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(i, off);
        const int ot = __shfl_xor(tag, off);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; tag = ot; }
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = i; st[w] = tag; }
    __syncthreads();
    v = sv[0]; i = si[0]; tag = st[0];
#pragma unroll
    for (int k = 1; k < PROWS / 64; ++k) {
        const double ov = sv[k];
        const int oi = si[k];
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; tag = st[k]; }
    }
}

template <class T>
__global__ __launch_bounds__(PROWS) void k_panel(int m, int n, T* __restrict__ A, int lda, int* __restrict__ ipiv,
                                                 int roff, char* __restrict__ ws, unsigned int bar_base) {
    __shared__ T as[LEAF][PROWS];   // the work-group's rows of the leaf, one row per thread
    __shared__ T urow[LEAF];
    __shared__ double sv[PROWS / 64];
    __shared__ int si[PROWS / 64], stg[PROWS / 64];
    unsigned int* bar = reinterpret_cast<unsigned int*>(ws + PanelWs::BAR);
    unsigned int* err = reinterpret_cast<unsigned int*>(ws + PanelWs::ERR);
    int* zcol = reinterpret_cast<int*>(ws + PanelWs::ZCOL);
    double* cval = reinterpret_cast<double*>(ws + PanelWs::CVAL);
    int* cidx = reinterpret_cast<int*>(ws + PanelWs::CIDX);
    double* crow = reinterpret_cast<double*>(ws + PanelWs::CROW);
    double* drow = reinterpret_cast<double*>(ws + PanelWs::DROW);

    const int P = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
    const int r = b * PROWS + tid;
    const bool own = r < m;
    for (int j = 0; j < n; ++j) as[j][tid] = own ? A[(size_t)j * lda + r] : T(0);

    for (int j = 0; j < n; ++j) {
        const int par = j & 1;
        double v = (own && r >= j) ? fabs((double)as[j][tid]) : -1.0;
        int idx = own ? r : 0x7fffffff;
        int tag = 0;
        block_argmax(v, idx, tag, sv, si, stg);
        if (tid == 0) {
            st_relaxed(&cval[par * MAXP + b], v);
            st_relaxed(&cidx[par * MAXP + b], idx);
        }
        if (own && r == idx && v >= 0.0)
            for (int jj = 0; jj < n; ++jj) st_relaxed(&crow[((size_t)par * MAXP + b) * LEAF + jj], (double)as[jj][tid]);
        if (own && r == j)
            for (int jj = 0; jj < n; ++jj) st_relaxed(&drow[par * LEAF + jj], (double)as[jj][tid]);
        grid_barrier(bar, bar_base + (unsigned)(j + 1) * (unsigned)P, err);
        double gv = -2.0;
        int gi = 0x7fffffff, gb = 0;
        if (tid < P) { gv = ld_relaxed(&cval[par * MAXP + tid]); gi = ld_relaxed(&cidx[par * MAXP + tid]); gb = tid; }
        block_argmax(gv, gi, gb, sv, si, stg);
        if (tid < n) urow[tid] = (T)ld_relaxed(&crow[((size_t)par * MAXP + gb) * LEAF + tid]);
        __syncthreads();
        const int p = gi;
        if (b == 0 && tid == 0) ipiv[j] = p + roff;
        if (own) {
            if (r == j) {
                for (int jj = 0; jj < n; ++jj) as[jj][tid] = urow[jj];
            } else if (r == p) {
                for (int jj = 0; jj < n; ++jj) as[jj][tid] = (T)ld_relaxed(&drow[par * LEAF + jj]);
            }
        }
        const T piv = urow[j];
        if (piv == T(0)) {
            if (b == 0 && tid == 0) {
                __hip_atomic_fetch_or(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicCAS(zcol, 0, roff + j + 1);
            }
        } else if (own && r > j) {
            const T l = as[j][tid] / piv;
            as[j][tid] = l;
            for (int jj = j + 1; jj < n; ++jj) as[jj][tid] -= l * urow[jj];
        }
        __syncthreads();
    }
    if (own)
        for (int j = 0; j < n; ++j) A[(size_t)j * lda + r] = as[j][tid];
}

// row interchanges k1..k2-1 (absolute rows ipiv[k] - roff relative to A) on ncols columns
template <class T>
__global__ __launch_bounds__(256) void k_laswp(int ncols, T* __restrict__ A, int lda, const int* __restrict__ ipiv,
                                               int k1, int k2, int roff) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    T* col = A + (size_t)c * lda;
    for (int k = k1; k < k2; ++k) {
        const int p = ipiv[k] - roff;
        if (p != k) { const T t = col[k]; col[k] = col[p]; col[p] = t; }
    }
}

// One 32-column step of the interchange-free factorisation of a panel's top
// block.  The block column starts at A (its diagonal element) and has m rows
// and bs <= 32 columns; U12 is the ncols columns to its right (rows 0..bs-1).
// Every work-group factors the bs x bs diagonal block in wave 0 (a row per
// lane, in registers, the pivot row read with v_readlane) and keeps it in
// LDS; then work-groups [0, prow) eliminate 256 rows each below the diagonal
// block (a row per thread, in registers) and the others solve 256 columns
// each of U12 = L11^-1 A12 (a column per thread), with no inter-work-group
// synchronisation at all.  Multipliers are formed with the reciprocal of the
// pivot, as LAPACK's getf2 does.  max |l| over all multipliers goes to *maxl
// (bits of a non-negative double; NaN sorts above every number).  If
// max |l| <= 1, partial pivoting with first-maximum ties would have chosen
// the diagonal at every column, i.e. this is exactly its elimination.  The
// factored diagonal block goes to dscr (NPB x NPB, column-major), never into
// A, because other work-groups may still be reading A's original one; the
// next launch on the stream (k_np_update or k_np_diag_store) moves it there.
constexpr int NPB = 32;

__device__ inline float lane_value(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double lane_value(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned int)(b & 0xffffffffll), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

template <class T>
__global__ __launch_bounds__(256) void k_np_step(int m, int bs, int ncols, int prow, T* __restrict__ A, int lda,
                                                 unsigned long long* __restrict__ maxl, T* __restrict__ dscr) {
    __shared__ T top[NPB][NPB + 1];
    __shared__ T invd[NPB];
    __shared__ unsigned long long red[4];
    const int tid = threadIdx.x, lane = tid & 63;
    for (int e = tid; e < bs * bs; e += 256) {
        const int i = e % bs, j = e / bs;
        top[i][j] = A[(size_t)j * lda + i];
    }
    __syncthreads();
    double lmax = 0.0;
    if (tid < 64) {
        const bool row = lane < bs;
        T r[NPB];
#pragma unroll
        for (int k = 0; k < NPB; ++k) r[k] = (row && k < bs) ? top[lane][k] : T(0);
#pragma unroll
        for (int j = 0; j < NPB; ++j) {
            if (j < bs) {
                const T inv = T(1) / lane_value(r[j], j);
                const bool below = row && lane > j;
                const T l = r[j] * inv;
                if (below) {
                    r[j] = l;
                    const double a = fabs((double)l);
                    if (blockIdx.x == 0 && !(a <= lmax)) lmax = a;   // reported once; NaN sticks
                }
#pragma unroll
                for (int k = j + 1; k < NPB; ++k) {
                    const T pk = lane_value(r[k], j);
                    if (below) r[k] -= l * pk;
                }
                if (lane == j) invd[j] = inv;
            }
        }
#pragma unroll
        for (int k = 0; k < NPB; ++k)
            if (row && k < bs) top[lane][k] = r[k];
    }
    __syncthreads();
    if ((int)blockIdx.x < prow) {
        // rows below the diagonal block
        const int r = bs + blockIdx.x * 256 + tid;
        if (r < m) {
            T a[NPB];
#pragma unroll
            for (int j = 0; j < NPB; ++j) a[j] = j < bs ? A[(size_t)j * lda + r] : T(0);
#pragma unroll
            for (int j = 0; j < NPB; ++j) {
                if (j < bs) {
                    const T l = a[j] * invd[j];
                    a[j] = l;
                    const double al = fabs((double)l);
                    if (!(al <= lmax)) lmax = al;
#pragma unroll
                    for (int k = j + 1; k < NPB; ++k) a[k] -= l * top[j][k];   // k >= bs: never stored
                }
            }
#pragma unroll
            for (int j = 0; j < NPB; ++j)
                if (j < bs) A[(size_t)j * lda + r] = a[j];
        }
    } else {
        // U12 = L11^-1 A12
        const int c = (blockIdx.x - prow) * 256 + tid;
        if (c < ncols) {
            T* col = A + (size_t)(bs + c) * lda;
            T u[NPB];
#pragma unroll
            for (int i = 0; i < NPB; ++i) u[i] = i < bs ? col[i] : T(0);
#pragma unroll
            for (int k = 0; k < NPB; ++k) {   // column-oriented forward substitution; i >= bs: never stored
                if (k < bs) {
#pragma unroll
                    for (int i = k + 1; i < NPB; ++i) u[i] -= top[i][k] * u[k];
                }
            }
#pragma unroll
            for (int i = 0; i < NPB; ++i)
                if (i < bs) col[i] = u[i];
        }
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(lmax);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if (lane == 0) red[tid >> 6] = bits;
    __syncthreads();
    if (tid == 0) {
        unsigned long long b = red[0];
        for (int w = 1; w < 4; ++w) b = red[w] > b ? red[w] : b;
        atomicMax(maxl, b);
    }
    if (blockIdx.x == 0)
        for (int e = tid; e < bs * bs; e += 256) {
            const int i = e % bs, j = e / bs;
            dscr[j * NPB + i] = top[i][j];
        }
}

// One launch factors a panel's whole top block (n x n) without interchanges:
// one work-group per 32-column stripe and one thread per row of the block.
// Stripes are taken in dispatch order through a ticket, so every stripe a
// work-group waits for belongs to a work-group that is already running (no
// co-residency assumption).  Work-group s applies the updates of stripes
// 0..s-1 in order as they are published: the 32 U rows of its stripe by
// forward substitution with stripe j's unit-lower diagonal block, then the
// rank-32 update of its rows below them; then it factors its own diagonal
// block (one wave, rows in registers, pivot row by v_readlane, reciprocal
// pivots as getf2) and the rows below it, merges max |l| into sync[0..1] and
// publishes the stripe (flag = epoch).  Published values go through
// agent-scope atomic stores and loads (coherent across the chip's L2 caches)
// with a release / acquire pair on the flag.  A waiting work-group gives up
// after ~0.2 s and raises sync[3], so the host reports an error instead of
// hanging.  One launch per panel replaces 2 n / 32 dependent launches, each
// of which had to find compute-unit slots behind the concurrent
// trailing-update GEMM.
constexpr int TOP_STRIPES = 64;
constexpr int TOP_SYNC_WORDS = 16 + TOP_STRIPES;   // [0,1] max |l| bits, [2] ticket, [3] error, [16..] flags
template <class T> constexpr int top_max_rows() { return sizeof(T) == 8 ? 768 : 1024; }
// FP32: 32-row blocks on the matrix cores, up to TOP_RBW per wave (8 waves at
// most, two per SIMD, so that a thread may hold 256 registers); FP64: a thread
// per row.  (Giving wave 0 the diagonal row block alone, so that it could
// factor it while the others still update, measured slower: sharing its SIMD
// with an updating wave doubled the factorisation's latency.)
constexpr int TOP_RBW = 4;
template <class T> constexpr int top_max_threads() { return sizeof(T) == 8 ? 768 : 512; }
template <class T> inline int top_threads(int n) {
    if (sizeof(T) == 8) return std::max(64, (n + 63) / 64 * 64);
    return 64 * std::max(1, ((n + NPB - 1) / NPB + TOP_RBW - 1) / TOP_RBW);
}
static_assert((1024 / NPB + TOP_RBW - 1) / TOP_RBW <= 8, "FP32 top block of 1024 rows needs more than 8 waves");

template <class T>
__device__ inline T ld_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// An SGPR zero the compiler can neither fold nor hoist: LDS tables read
// through `base + zero` keep one base register, where constant addresses are
// materialised one register each.
__device__ inline int opaque_zero() {
    int z;
    __asm__ volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}
// Ends one step of a long unrolled multiply-add sweep: the running values are
// pinned to registers here and LDS reads may not move across.  Without it the
// compiler issues all the sweep's LDS reads first and sinks the arithmetic to
// the end, spilling hundreds of registers.
template <class T, int N>
__device__ inline void step_fence(T (&v)[N]) {
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < N; ++i) __asm__ volatile("" : "+v"(v[i]));
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

// FP32 32-row blocks on the matrix cores: acc (32 x 32, v_mfma_f32_32x32x2f32
// layout) += W^T X^T over k = 0..31, i.e. acc[c][i] = sum_k W[k][c] X[i][k],
// where the lane of row i (i = lane % 32) supplies X[i][k] for k = 2 kk +
// lane / 32 in xv[kk] and W is a 32 x 32 LDS table.  Lane l, register q of the
// result holds column c = 8 (q / 4) + 4 (l / 32) + q % 4 of row i.
__device__ inline floatx16 mfma_rows32(const float* __restrict__ W, const float (&xv)[16], int lane) {
    floatx16 acc = {};
    const int c = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(W[(2 * kk + kh) * NPB + c], xv[kk], acc, 0, 0, 0);
    return acc;
}

// Unit-lower inverse (lane = column c < 32 of L^-1) of the 32 x 32 LU block
// in Lb (row stride NPB + 1), written transposed to W[c * NPB + i] =
// (L^-1)[i][c] through `st`.  Column oriented: once x[k] is final, step k
// subtracts L[i][k] x[k] from every x[i] below it, so the 31 - k updates of a
// step are independent and the chain from step to step is one FMA (each x[i]
// still receives its terms in increasing k, as forward substitution adds
// them).  The fence after each step keeps the later steps' LDS reads from
// piling up in registers.
template <class T, class St>
__device__ inline void unit_lower_inverse(const T* __restrict__ Lb, int c, St st) {
    T x[NPB];
#pragma unroll
    for (int i = 0; i < NPB; ++i) x[i] = i == c ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < NPB - 1; ++k) {
        const T xk = x[k];
#pragma unroll
        for (int i = k + 1; i < NPB; ++i) x[i] -= Lb[i * (NPB + 1) + k] * xk;
        step_fence(x);
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) st(c * NPB + i, x[i]);
}
// Upper inverse (lane = column c < 32 of U^-1, with the reciprocal pivots),
// written to W[i * NPB + c] = (U^-1)[i][c].  Column oriented from the last
// column: x[k] = acc[k] / u_kk is final once the columns beyond k are applied,
// then U[i][k] x[k] leaves every acc[i] above it (independent updates).
template <class T>
__device__ inline void upper_inverse(const T* __restrict__ Lb, const T* __restrict__ invd, int c, T* __restrict__ W) {
    T x[NPB];
#pragma unroll
    for (int i = 0; i < NPB; ++i) x[i] = i == c ? T(1) : T(0);
#pragma unroll
    for (int k = NPB - 1; k >= 0; --k) {
        x[k] *= invd[k];
        const T xk = x[k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] -= Lb[i * (NPB + 1) + k] * xk;
        step_fence(x);
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) W[i * NPB + c] = x[i];
}

// The same two inverses from the block stored by columns (LT[k * NPB + i] =
// entry (i, k), 16-byte aligned, in LDS): a step reads its whole column with
// eight 16-byte loads and waits for them once, not for one load per FMA.
__device__ inline void lds_col(const float* LT, int k, float (&v)[NPB]) {
    const float4* p = reinterpret_cast<const float4*>(LT + k * NPB);
#pragma unroll
    for (int q = 0; q < NPB / 4; ++q) {
        const float4 t = p[q];
        v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
    }
}
// (LT is not __restrict__: the fences must keep its reads in their steps;
// x[k] is pinned once final, or the compiler keeps the steps' products
// apart and spills)
template <class St>
__device__ inline void unit_lower_inverse_cols(const float* LT, int c, St st) {
    float x[NPB];
#pragma unroll
    for (int i = 0; i < NPB; ++i) x[i] = i == c ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < NPB - 1; ++k) {
        float cur[NPB];   // column k
        lds_col(LT, k, cur);
        step_fence(cur);
        const float xk = x[k];
#pragma unroll
        for (int i = k + 1; i < NPB; ++i) x[i] -= cur[i] * xk;
        __asm__ volatile("" : "+v"(x[k + 1]));
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) st(c * NPB + i, x[i]);
}
__device__ inline void upper_inverse_cols(const float* LT, const float* __restrict__ invd, int c,
                                          float* __restrict__ W) {
    float x[NPB];
#pragma unroll
    for (int i = 0; i < NPB; ++i) x[i] = i == c ? 1.0f : 0.0f;
#pragma unroll
    for (int k = NPB - 1; k >= 0; --k) {
        float cur[NPB];   // column k
        lds_col(LT, k, cur);
        step_fence(cur);
        x[k] *= invd[k];
        __asm__ volatile("" : "+v"(x[k]));
        const float xk = x[k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] -= cur[i] * xk;
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) W[i * NPB + c] = x[i];
}

// FP32 row block in slot h of wave wv (nw waves), or -1 past the block
__device__ inline int top_rb(int wv, int h, int nw, int n) {
    const int rb = wv + h * nw;
    return rb * NPB < n ? rb : -1;
}

template <class T>
__global__ __launch_bounds__(top_max_threads<T>()) void k_np_top(int n, T* __restrict__ A, int lda,
                                                                  unsigned int* __restrict__ sync,
                                                                  unsigned int epoch, T* __restrict__ linv) {
    constexpr int MAXW = top_max_threads<T>() / 64;
    // FP32 runs its block products on the matrix cores (32-row blocks, up to
    // TOP_RBW per wave) and publishes L^-1 of every diagonal block, so that a
    // stripe's U rows are one product too; FP64 keeps one row per thread on
    // the vector units and forward substitution
    constexpr bool MF = sizeof(T) == 4;
    __shared__ T Ld[NPB][NPB + 1];   // stripe j's unit-lower diagonal block (FP64) / own factored block
    __shared__ T Us[NPB][NPB];       // this stripe's 32 U rows of step j
    __shared__ T Wt[NPB * NPB];      // FP32: published L^-1 of stripe j (transposed) / own U^-1
    __shared__ T invd[NPB];
    __shared__ unsigned long long red[MAXW];
    __shared__ int s_stripe, s_go;
    unsigned int* ticket = sync + 2;
    unsigned int* err = sync + 3;
    unsigned int* flags = sync + 16;
    const int tid = threadIdx.x, lane = tid & 63, nt = blockDim.x;
    const int wv = tid >> 6, nw = nt >> 6;
    if (tid == 0) s_stripe = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int s = s_stripe;
    const int c0 = s * NPB, w = min(NPB, n - c0);
    T* S = A + (size_t)c0 * lda;   // this stripe: columns c0 .. c0 + w - 1, all n rows
    const int r = tid;              // FP64: this thread's row
    // max |l| as the bits of a non-negative double: such doubles (and NaN
    // above them) order as unsigned integers; exact for FP32 and FP64 alike
    unsigned long long lbits = 0;
    for (int j = 0; j < s; ++j) {
        const int r0 = j * NPB;
        if (tid == 0) {
            int go = 1;
            unsigned int spins = 0;
            while (__hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 22) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    go = 0;
                    break;
                }
            }
            // one acquire for the work-group (the barrier below extends it)
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;   // error raised: results invalid, every work-group leaves
        const T* Lj = A + (size_t)r0 * lda;   // stripe j (published)
        // FP32: the multipliers of stripe j for this wave's row blocks, loaded
        // now so that their latency overlaps the U rows
        float xb[TOP_RBW][16];
        if constexpr (MF) {
#pragma unroll
            for (int h = 0; h < TOP_RBW; ++h) {
                const int rb = top_rb(wv, h, nw, n), i = min(max(rb, 0) * NPB + (lane & 31), n - 1);
                const bool act = rb > j;
#pragma unroll
                for (int kk = 0; kk < 16; ++kk)
                    xb[h][kk] = act ? ld_agent(reinterpret_cast<const float*>(Lj) + (size_t)(2 * kk + (lane >> 5)) * lda + i)
                                    : 0.0f;
            }
        }
        for (int e = tid; e < NPB * NPB; e += nt) {
            const int i = e & (NPB - 1), k = e / NPB;
            if constexpr (MF) Wt[e] = ld_agent(linv + (size_t)j * NPB * NPB + e);
            else Ld[i][k] = ld_agent(&Lj[(size_t)k * lda + r0 + i]);
            Us[i][k] = k < w ? S[(size_t)k * lda + r0 + i] : T(0);
        }
        __syncthreads();
        // U rows r0.. of this stripe = L_jj^-1 A[r0.., stripe]
        if constexpr (MF) {
            if (wv == 0) {
                // acc[i][c] = sum_k (L^-1)[i][k] A[k][c]: lane c supplies A[k][c]
                float xv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) xv[kk] = reinterpret_cast<const float*>(&Us[0][0])[(2 * kk + (lane >> 5)) * NPB + (lane & 31)];
                const floatx16 acc = mfma_rows32(reinterpret_cast<const float*>(Wt), xv, lane);
                const int c = lane & 31;
                float* Sf = reinterpret_cast<float*>(S);
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int i = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    if (c < w) {
                        reinterpret_cast<float*>(&Us[0][0])[i * NPB + c] = acc[q];
                        Sf[(size_t)c * lda + r0 + i] = acc[q];
                    }
                }
            }
        } else if (tid < w) {
            // a column per lane of wave 0
            T u[NPB];
#pragma unroll
            for (int i = 0; i < NPB; ++i) u[i] = Us[i][tid];
            const T* Lb = &Ld[0][0] + opaque_zero();
#pragma unroll
            for (int k = 0; k < NPB; ++k) {
#pragma unroll
                for (int i = k + 1; i < NPB; ++i) u[i] -= Lb[i * (NPB + 1) + k] * u[k];
                step_fence(u);
            }
#pragma unroll
            for (int i = 0; i < NPB; ++i) {
                Us[i][tid] = u[i];
                S[(size_t)tid * lda + r0 + i] = u[i];
            }
        }
        __syncthreads();
        // rank-32 update of this stripe's rows below: a[r, :] -= L_j[r, :] U
        if constexpr (MF) {
#pragma unroll
            for (int h = 0; h < TOP_RBW; ++h) {
                const int rb = top_rb(wv, h, nw, n);
                if (rb > j) {
                    const floatx16 acc = mfma_rows32(reinterpret_cast<const float*>(&Us[0][0]), xb[h], lane);
                    const int i = rb * NPB + (lane & 31);
                    float* Sf = reinterpret_cast<float*>(S) + opaque_zero();   // not hoisted out of the step loop
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        const int c = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                        if (i < n && c < w) Sf[(size_t)c * lda + i] -= acc[q];
                    }
                }
                __asm__ volatile("" ::: "memory");   // one row block at a time (registers)
            }
        } else if (r >= r0 + NPB && r < n) {
            // opaque bases: the stripe's 32 column addresses would otherwise be
            // hoisted out of the step loop and held in 64 registers throughout
            const T* Ub = &Us[0][0] + opaque_zero();
            T l[NPB];
#pragma unroll
            for (int k = 0; k < NPB; ++k) l[k] = ld_agent(Lj + (size_t)k * lda + r);
            // in two passes of 16 columns (registers)
            constexpr int CB = 16;
#pragma unroll
            for (int cb = 0; cb < NPB; cb += CB) {
                T* So = S + opaque_zero();
                T a[CB];
#pragma unroll
                for (int q = 0; q < CB; ++q) {
                    const T v = (So + (size_t)min(cb + q, w - 1) * lda)[r];   // clamped: in bounds for a narrow stripe
                    a[q] = cb + q < w ? v : T(0);
                }
#pragma unroll
                for (int k = 0; k < NPB; ++k) {
#pragma unroll
                    for (int q = 0; q < CB; ++q) a[q] -= l[k] * Ub[k * NPB + cb + q];
                    step_fence(a);
                }
#pragma unroll
                for (int q = 0; q < CB; ++q)
                    if (cb + q < w) (So + (size_t)(cb + q) * lda)[r] = a[q];
            }
        }
        __syncthreads();   // Ld / Us / Wt are refilled by the next step
    }
    // own diagonal block (rows c0 .., w x w), factored by wave 0 in registers
    // (pivot row by v_readlane); rows at or above the pivot get a zero
    // multiplier instead of a branch.  Ld and invd are zero beyond w.
    if (tid < 64) {
        const bool row = lane < w;
        T rw[NPB];
        unsigned int fbits = 0;   // FP32: max |l| as float bits (cheaper on the elimination's chain)
#pragma unroll
        for (int k = 0; k < NPB; ++k) rw[k] = (row && k < w) ? S[(size_t)k * lda + c0 + lane] : T(0);
#pragma unroll
        for (int j = 0; j < NPB; ++j) {
            if (j < w) {
                const T piv = lane_value(rw[j], j);
                T inv;
                if constexpr (MF) inv = __builtin_amdgcn_rcpf(piv);
                else inv = T(1) / piv;
                const bool below = row && lane > j;
                const T l = below ? rw[j] * inv : T(0);
                if (below) rw[j] = l;
                if constexpr (MF) {
                    const unsigned int lb = __float_as_uint(fabsf(l));
                    fbits = lb > fbits ? lb : fbits;
                } else {
                    const unsigned long long lb = (unsigned long long)__double_as_longlong(fabs((double)l));
                    lbits = lb > lbits ? lb : lbits;
                }
                T pk[NPB];
#pragma unroll
                for (int k = j + 1; k < NPB; ++k) pk[k] = lane_value(rw[k], j);
#pragma unroll
                for (int k = j + 1; k < NPB; ++k) rw[k] -= l * pk[k];
                if (lane == j) invd[j] = inv;
            } else if (lane == j) {
                invd[j] = T(0);
            }
        }
        if (lane < NPB) {
#pragma unroll
            for (int k = 0; k < NPB; ++k) {
                Ld[lane][k] = (row && k < w) ? rw[k] : T(0);
                if (row && k < w) st_agent(&S[(size_t)k * lda + c0 + lane], rw[k]);
            }
        }
        if constexpr (MF) {
            const unsigned long long fd = fbits > 0x7F800000u ? 0x7FF8000000000000ull
                                                              : (unsigned long long)__double_as_longlong((double)__uint_as_float(fbits));
            lbits = fd > lbits ? fd : lbits;
        }
    }
    __syncthreads();
    if constexpr (MF) {
        // L^-1 of the diagonal block (published for the later stripes' U rows)
        // and U^-1 (for the rows below, L = A U^-1), by two waves at once
        const bool below = c0 + NPB < n;
        const float* Lb = reinterpret_cast<const float*>(&Ld[0][0]) + opaque_zero();
        if (wv == (1 % nw) && lane < NPB && w == NPB)
            unit_lower_inverse(Lb, lane, [&](int e, float v) { st_agent(reinterpret_cast<float*>(linv) + (size_t)s * NPB * NPB + e, v); });
        if (wv == (2 % nw) && lane < NPB && below)
            upper_inverse(Lb, reinterpret_cast<const float*>(invd), lane, reinterpret_cast<float*>(Wt));
        __syncthreads();
        // rows below the diagonal block on the matrix cores
        float* Sf = reinterpret_cast<float*>(S);
        unsigned int fb = 0;   // bits of max |l| as a float (non-negative floats and NaN order as uints)
#pragma unroll
        for (int h = 0; h < TOP_RBW; ++h) {
            const int rb = top_rb(wv, h, nw, n);
            if (below && rb > s) {
                const int i = rb * NPB + (lane & 31), ic = min(i, n - 1);
                float xv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) xv[kk] = Sf[(size_t)(2 * kk + (lane >> 5)) * lda + ic];
                const floatx16 acc = mfma_rows32(reinterpret_cast<const float*>(Wt), xv, lane);
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int c = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    if (i < n) {
                        const float lq = acc[q];
                        st_agent(Sf + (size_t)c * lda + i, lq);
                        const unsigned int lb = __float_as_uint(fabsf(lq));
                        fb = lb > fb ? lb : fb;
                    }
                }
            }
        }
        const unsigned long long fbd = fb > 0x7F800000u ? 0x7FF8000000000000ull
                                                        : (unsigned long long)__double_as_longlong((double)__uint_as_float(fb));
        lbits = fbd > lbits ? fbd : lbits;
    } else if (r >= c0 + w && r < n) {
        // rows below the diagonal block, a row per thread
        const T* Lb = &Ld[0][0] + opaque_zero();
        T a[NPB];
#pragma unroll
        for (int j = 0; j < NPB; ++j) {
            const T v = (S + (size_t)min(j, w - 1) * lda)[r];
            a[j] = j < w ? v : T(0);
        }
#pragma unroll
        for (int j = 0; j < NPB; ++j) {
            const T l = a[j] * invd[j];
            a[j] = l;
            const unsigned long long lb = (unsigned long long)__double_as_longlong(fabs((double)l));
            lbits = lb > lbits ? lb : lbits;
#pragma unroll
            for (int k = j + 1; k < NPB; ++k) a[k] -= l * Lb[j * (NPB + 1) + k];
            step_fence(a);
        }
#pragma unroll
        for (int j = 0; j < NPB; ++j)
            if (j < w) st_agent(S + (size_t)j * lda + r, a[j]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(lbits, off);
        lbits = o > lbits ? o : lbits;
    }
    if (lane == 0) red[wv] = lbits;
    __syncthreads();
    if (tid == 0) {
        unsigned long long bits = red[0];
        for (int q = 1; q < nw; ++q) bits = red[q] > bits ? red[q] : bits;
        atomicMax(reinterpret_cast<unsigned long long*>(sync), bits);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[s], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// k order of a 32x32x2 product whose row operand comes straight from an
// accumulator (lane half kh holds columns 8 g + 4 kh + e, g, e < 4): step kk
// covers k = 8 (kk / 4) + 4 kh + kk % 4, the same k for both operands
__device__ inline int kperm(int kk, int kh) { return 8 * (kk >> 2) + 4 * kh + (kk & 3); }
__device__ inline floatx16 mfma_rows32p(const float* __restrict__ W, const floatx16& xv, int lane) {
    floatx16 acc = {};
    const int c = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(W[kperm(kk, kh) * NPB + c], xv[kk], acc, 0, 0, 0);
    return acc;
}

// FP32 top block, the same stripe algorithm as k_np_top, with every work-group's
// stripe held in registers for the whole launch (its row blocks in the
// accumulator layout of the 32x32 products: lane row rb * 32 + lane % 32,
// register q column 8 (q / 4) + 4 (lane / 32) + q % 4): a step j only reads
// stripe j's multipliers and L^-1 (published by its work-group) and updates
// registers, where k_np_top read and wrote the whole stripe in memory every
// step.  A row block that another wave needs (the step's own row block, the
// diagonal block) goes through LDS.  Only final values are stored: the U rows
// of every step, the factored diagonal block and the multipliers below it.
// Stripe j publishes under two flags: its L^-1 and the multipliers of row
// block j + 1 (all that stripe j + 1 needs before factoring its diagonal
// block), then its other multipliers.  Stripe j + 1 applies step j to its
// diagonal row block alone, factors it, and applies step j to its other row
// blocks while that runs; the stripes further right apply step j to their
// row blocks below j + 1 after the second flag.
// backup (optional, ld n) receives the block's original values.
// Batched (gticket set): one launch factors the top blocks of a batch of
// problems, `stripes` stripes each; the ticket taken from gticket names the
// problem (ticket / stripes) and the stripe, and problem b's block, sync words
// and published inverses start b sA, b sstride, b slinv elements further.
// Tickets are dealt in dispatch order, so every stripe waited for belongs to
// the same problem and to an already running work-group.
__global__ __launch_bounds__(512) void k_np_top32(int n, float* __restrict__ A, int lda,
                                                  unsigned int* __restrict__ sync, unsigned int epoch,
                                                  float* __restrict__ linv, float* __restrict__ backup,
                                                  unsigned int* __restrict__ gticket, int stripes, long long sA,
                                                  int sstride, long long slinv,
                                                  const unsigned int* __restrict__ ebase) {
    // the publication stamp: the host's count of launches, plus a device base
    // that the batched path advances before each factorisation (so launches
    // replayed from a graph stamp new values)
    if (ebase) epoch += *ebase;
    constexpr int MAXW = 8;
    __shared__ float Us[NPB * NPB];      // U rows of step j: Us[i * NPB + c]
    __shared__ float Wt[NPB * NPB];      // published L^-1 of stripe j (transposed) / own U^-1
    __shared__ float Xs[NPB][NPB + 1];   // a row block of this stripe handed between waves: Xs[i][c]
    __shared__ float Ld[NPB][NPB + 1];   // own factored diagonal block
    __shared__ __align__(16) float LdT[NPB * NPB];   // the same by columns, for the inverses
    __shared__ float invd[NPB];
    __shared__ unsigned long long red[MAXW];
    __shared__ int s_stripe, s_go;
    const int tid = threadIdx.x, lane = tid & 63, nt = blockDim.x;
    const int wv = tid >> 6, nw = nt >> 6;
    const int li = lane & 31, kh = lane >> 5;
    if (tid == 0) s_stripe = (int)atomicAdd(gticket ? gticket : sync + 2, 1u);
    __syncthreads();
    int s = s_stripe;
    if (gticket) {
        const int q = s / stripes;
        s -= q * stripes;
        A += q * sA;
        sync += (long long)q * sstride;
        linv += q * slinv;
    }
    unsigned int* err = sync + 3;
    // flags[j]: stripe j's L^-1 and the multipliers of row block j + 1 are
    // published; flags[32 + j]: all of its multipliers are
    unsigned int* flags = sync + 16;
    const int c0 = s * NPB, w = min(NPB, n - c0);
    float* S = A + (size_t)c0 * lda;
    floatx16 sv[TOP_RBW];
#pragma unroll
    for (int h = 0; h < TOP_RBW; ++h) {
        const int rb = top_rb(wv, h, nw, n), i = rb * NPB + li;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = 8 * (q >> 2) + 4 * kh + (q & 3);
            float v = 0.0f;
            if (rb >= 0 && i < n && c < w) {
                v = S[(size_t)c * lda + i];
                if (backup) backup[(size_t)(c0 + c) * n + i] = v;
            }
            sv[h][q] = v;
        }
    }
    // the wave holding row block rb writes it to Xs
    auto hand_over = [&](int rb) {
        if (wv == rb % nw) {
            const int h = rb / nw;
#pragma unroll
            for (int hh = 0; hh < TOP_RBW; ++hh)
                if (hh == h)
#pragma unroll
                    for (int q = 0; q < 16; ++q) Xs[li][8 * (q >> 2) + 4 * kh + (q & 3)] = sv[hh][q];
        }
    };
    unsigned long long lbits = 0;
    // wait for flag f of this launch (lane 0 spins, bounded; false: the wait
    // gave up or another work-group did, the error word is raised)
    auto wait_flag = [&](const unsigned int* fl) {
        int go = 1;
        unsigned int spins = 0;
        while (__hip_atomic_load(fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 22) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                go = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return go;
    };
    for (int j = 0; j < s; ++j) {
        const int r0 = j * NPB;
        if (tid == 0) s_go = wait_flag(&flags[j]);
        __syncthreads();
        if (!s_go) return;   // error raised: results invalid, every work-group leaves
        const float* Lj = A + (size_t)r0 * lda;   // stripe j (published)
        // first row block j + 1 alone (its multipliers come with the first
        // flag); the step before this stripe's own leaves the rows below its
        // diagonal block for while that block is factored, the others apply
        // the step to them after the second flag
        const bool last = j == s - 1;
        // the slot of row block j + 1 in this wave (-1: another wave holds it)
        int h1 = -1;
#pragma unroll
        for (int h = 0; h < TOP_RBW; ++h)
            if (top_rb(wv, h, nw, n) == j + 1) h1 = h;
        float xb[16];
        {
            const int i = min((j + 1) * NPB + li, n - 1);
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) xb[kk] = h1 >= 0 ? ld_agent(Lj + (size_t)(2 * kk + kh) * lda + i) : 0.0f;
        }
        for (int e = tid; e < NPB * NPB; e += nt) Wt[e] = ld_agent(linv + (size_t)j * NPB * NPB + e);
        hand_over(j);
        __syncthreads();
        // U rows r0.. of this stripe = L_jj^-1 A[r0.., stripe] (final values)
        if (wv == 0) {
            float xv[16];
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) xv[kk] = Xs[2 * kk + kh][li];
            const floatx16 acc = mfma_rows32(Wt, xv, lane);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int i = 8 * (q >> 2) + 4 * kh + (q & 3);
                Us[i * NPB + li] = acc[q];
                if (li < w) S[(size_t)li * lda + r0 + i] = acc[q];
            }
        }
        __syncthreads();
        // rank-32 update of this stripe's row blocks below, in registers:
        // row block j + 1 now
        if (h1 >= 0) {
            const floatx16 acc = mfma_rows32(Us, xb, lane);
#pragma unroll
            for (int h = 0; h < TOP_RBW; ++h)
                if (h == h1) sv[h] -= acc;
        }
        if (last) break;   // the last step's Us stays for the deferred rows
        // the rows below it once stripe j's other multipliers are published
        if (tid == 0) s_go = wait_flag(&flags[32 + j]);
        __syncthreads();
        if (!s_go) return;
#pragma unroll
        for (int h = 0; h < TOP_RBW; ++h) {
            const int rb = top_rb(wv, h, nw, n);
            if (rb > j + 1) {
                const int i = min(rb * NPB + li, n - 1);
                float xv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) xv[kk] = ld_agent(Lj + (size_t)(2 * kk + kh) * lda + i);
                const floatx16 acc = mfma_rows32(Us, xv, lane);
                sv[h] -= acc;
            }
        }
        __syncthreads();   // Us / Wt / Xs are refilled by the next step
    }
    // the deferred rows: the last step (stripe s - 1, its U rows still in Us)
    // applied to this wave's row blocks below the diagonal one, multipliers
    // read again (they were published before that step's flag)
    auto deferred = [&] {
        if (s == 0) return;
        // stripe s - 1's multipliers below row block s (this wave waits for
        // them itself; a wait that gives up leaves the rows as they are, the
        // error word is raised)
        if (!__builtin_amdgcn_readfirstlane(lane == 0 ? wait_flag(&flags[32 + s - 1]) : 0)) return;
        const float* Lj = A + (size_t)(s - 1) * NPB * lda;
#pragma unroll
        for (int h = 0; h < TOP_RBW; ++h) {
            const int rb = top_rb(wv, h, nw, n);
            if (rb > s) {
                const int i = min(rb * NPB + li, n - 1);
                float xv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) xv[kk] = ld_agent(Lj + (size_t)(2 * kk + kh) * lda + i);
                const floatx16 acc = mfma_rows32(Us, xv, lane);
                sv[h] -= acc;
            }
        }
    };
    // own diagonal block: its row block from the holding wave, factored by
    // wave 0 in registers (pivot row by v_readlane); meanwhile the waves on
    // the other three SIMDs (waves are dealt to SIMDs round robin) take the
    // deferred rows, the ones on wave 0's SIMD (and wave 0) during the
    // inverses below
    hand_over(s);
    __syncthreads();
    if (tid >= 64 && (wv & 3) != 0) deferred();
    if (tid < 64) {
        const bool row = lane < w;
        float rw[NPB];
        unsigned int fbits = 0;
        float myinv = 0.0f;   // lane j: the reciprocal of pivot j (written to LDS once at the end)
#pragma unroll
        for (int k = 0; k < NPB; ++k) rw[k] = (row && k < w) ? Xs[lane & 31][k] : 0.0f;
        auto pivot_step = [&](int j) {
            const float piv = lane_value(rw[j], j);
            const float inv = __builtin_amdgcn_rcpf(piv);
            const bool below = row && lane > j;
            const float l = below ? rw[j] * inv : 0.0f;
            if (below) rw[j] = l;
            const unsigned int lb = __float_as_uint(fabsf(l));
            fbits = lb > fbits ? lb : fbits;
            float pk[NPB];
#pragma unroll
            for (int k = j + 1; k < NPB; ++k) pk[k] = lane_value(rw[k], j);
#pragma unroll
            for (int k = j + 1; k < NPB; ++k) rw[k] -= l * pk[k];
            myinv = lane == j ? inv : myinv;
        };
        if (w == NPB) {   // a full stripe: no per-step test
#pragma unroll
            for (int j = 0; j < NPB; ++j) pivot_step(j);
        } else {
#pragma unroll
            for (int j = 0; j < NPB; ++j)
                if (j < w) pivot_step(j);
        }
        if (lane < NPB) {
            invd[lane] = myinv;
#pragma unroll
            for (int k = 0; k < NPB; ++k) {
                const float v = (row && k < w) ? rw[k] : 0.0f;
                Ld[lane][k] = v;
                LdT[k * NPB + lane] = v;
                if (row && k < w) st_agent(&S[(size_t)k * lda + c0 + lane], rw[k]);
            }
        }
        const unsigned long long fd = fbits > 0x7F800000u ? 0x7FF8000000000000ull
                                                          : (unsigned long long)__double_as_longlong((double)__uint_as_float(fbits));
        lbits = fd > lbits ? fd : lbits;
    }
    __syncthreads();
    const bool below = c0 + NPB < n;
    {
        const float* LT = LdT + opaque_zero();
        if (wv == (1 % nw) && lane < NPB && w == NPB)
            unit_lower_inverse_cols(LT, lane, [&](int e, float v) { st_agent(linv + (size_t)s * NPB * NPB + e, v); });
        if (wv == (2 % nw) && lane < NPB && below) upper_inverse_cols(LT, invd, lane, Wt);
    }
    if ((wv & 3) == 0) deferred();
    __syncthreads();
    // multipliers of the row blocks below: L = A U^-1, the row operand from
    // registers; row block s + 1 first, published with the first flag by
    // the wave that holds it (wave 0 when there is none)
    unsigned int fb = 0;
    auto mult = [&](int h) {
        const int rb = top_rb(wv, h, nw, n);
        const int i = rb * NPB + li;
        const floatx16 acc = mfma_rows32p(Wt, sv[h], lane);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = 8 * (q >> 2) + 4 * kh + (q & 3);
            if (i < n) {
                const float lq = acc[q];
                st_agent(S + (size_t)c * lda + i, lq);
                const unsigned int lb = __float_as_uint(fabsf(lq));
                fb = lb > fb ? lb : fb;
            }
        }
    };
#pragma unroll
    for (int h = 0; h < TOP_RBW; ++h)
        if (below && top_rb(wv, h, nw, n) == s + 1) mult(h);
    if (wv == (below ? (s + 1) % nw : 0) && lane == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[s], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int h = 0; h < TOP_RBW; ++h)
        if (below && top_rb(wv, h, nw, n) > s + 1) mult(h);
    const unsigned long long fbd = fb > 0x7F800000u ? 0x7FF8000000000000ull
                                                    : (unsigned long long)__double_as_longlong((double)__uint_as_float(fb));
    lbits = fbd > lbits ? fbd : lbits;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(lbits, off);
        lbits = o > lbits ? o : lbits;
    }
    if (lane == 0) red[wv] = lbits;
    __syncthreads();
    if (tid == 0) {
        unsigned long long bits = red[0];
        for (int q = 1; q < nw; ++q) bits = red[q] > bits ? red[q] : bits;
        atomicMax(reinterpret_cast<unsigned long long*>(sync), bits);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&flags[32 + s], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// the factored diagonal block from k_np_step's scratch into A
template <class T>
__device__ inline void np_diag_store(int bs, const T* __restrict__ dscr, T* __restrict__ Adiag, int lda) {
    for (int e = threadIdx.x; e < bs * bs; e += blockDim.x) {
        const int i = e % bs, j = e / bs;
        Adiag[(size_t)j * lda + i] = dscr[j * NPB + i];
    }
}
template <class T>
__global__ __launch_bounds__(256) void k_np_diag_store(int bs, const T* __restrict__ dscr, T* __restrict__ Adiag,
                                                       int lda) {
    np_diag_store(bs, dscr, Adiag, lda);
}

// C (mrows x ncols) -= L (mrows x bs) U (bs x ncols), column-major, bs <= 32:
// the rank-bs update that follows a k_np_step, one 64 x 64 tile per work-group.
// Work-group (0, 0) also stores the step's factored diagonal block.
template <class T>
__global__ __launch_bounds__(256) void k_np_update(int mrows, int ncols, int bs, const T* __restrict__ L,
                                                   const T* __restrict__ U, T* __restrict__ C, int lda,
                                                   const T* __restrict__ dscr, T* __restrict__ Adiag) {
    __shared__ T ls[64][NPB + 1];
    __shared__ T us[NPB][64 + 1];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0) np_diag_store(bs, dscr, Adiag, lda);
    const int ti = blockIdx.x * 64, tj = blockIdx.y * 64;
    for (int e = tid; e < 64 * NPB; e += 256) {
        const int i = e & 63, k = e >> 6;
        ls[i][k] = (ti + i < mrows && k < bs) ? L[(size_t)k * lda + ti + i] : T(0);
    }
    for (int e = tid; e < NPB * 64; e += 256) {
        const int k = e % NPB, j = e / NPB;
        us[k][j] = (tj + j < ncols && k < bs) ? U[(size_t)(tj + j) * lda + k] : T(0);
    }
    __syncthreads();
    const int ty = tid >> 4, tx = tid & 15;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
#pragma unroll 8
    for (int k = 0; k < NPB; ++k) {
        T lv[4], uv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            lv[q] = ls[ty * 4 + q][k];
            uv[q] = us[k][tx * 4 + q];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += lv[a] * uv[b];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = tj + tx * 4 + b;
        if (c >= ncols) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int r = ti + ty * 4 + a;
            if (r < mrows) C[(size_t)c * lda + r] -= acc[a][b];
        }
    }
}

__global__ void k_iota(int* p, int n, int base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = base + i;
}

__global__ void k_fill_u32(unsigned int* p, int n, unsigned int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ rocBLAS glue --
static void trsm_llnu(rocblas_handle h, int m, int n, const double* A, int lda, double* B, int ldb) {
    const double one = 1.0;
    BH_BLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                          rocblas_diagonal_unit, m, n, &one, A, lda, B, ldb));
}
static void trsm_llnu(rocblas_handle h, int m, int n, const float* A, int lda, float* B, int ldb) {
    const float one = 1.0f;
    BH_BLAS(rocblas_strsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                          rocblas_diagonal_unit, m, n, &one, A, lda, B, ldb));
}
static void gemm_minus(rocblas_handle h, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                       double* C, int ldc) {
    const double alpha = -1.0, beta = 1.0;
    BH_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B,
                          ldb, &beta, C, ldc));
}
static void gemm_minus(rocblas_handle h, int m, int n, int k, const float* A, int lda, const float* B, int ldb,
                       float* C, int ldc) {
    const float alpha = -1.0f, beta = 1.0f;
    BH_BLAS(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B,
                          ldb, &beta, C, ldc));
}


// dst = src for a column-major rows x cols block (coalesced; the runtime's
// rectangular copy measured 4-5x slower on these shapes)
template <class T>
__global__ __launch_bounds__(256) void k_copy2d(int rows, int cols, const T* __restrict__ src, int lds,
                                                T* __restrict__ dst, int ldd) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const int c0 = blockIdx.y * 16, c1 = min(cols, c0 + 16);
    for (int c = c0; c < c1; ++c) dst[(size_t)c * ldd + r] = src[(size_t)c * lds + r];
}
template <class T>
static void copy2d(hipStream_t st, int rows, int cols, const T* src, int lds, T* dst, int ldd) {
    if (rows <= 0 || cols <= 0) return;
    k_copy2d<T><<<dim3((rows + 255) / 256, (cols + 15) / 16), 256, 0, st>>>(rows, cols, src, lds, dst, ldd);
    BH_HIP(hipGetLastError());
}

// A stream with the rocBLAS handle bound to it.
struct Lane {
    hipStream_t s;
    rocblas_handle h;
};

// ------------------------------------------- triangular solves by inverse --
// Large blocked triangular solves as one GEMM with the explicit inverse of the
// (at most NB x NB) triangle (trinv.hip).  The inverse of a unit lower triangle
// with |l| <= 1 (partial pivoting) or of U11 of such a factorisation is well
// conditioned for these matrices, and the factor only preconditions the FP64
// refinement of the mixed solvers.
static void gemm_nn(rocblas_handle h, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc) {
    const double one = 1.0, zero = 0.0;
    BH_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero,
                          C, ldc));
}
static void gemm_nn(rocblas_handle h, int m, int n, int k, const float* A, int lda, const float* B, int ldb,
                    float* C, int ldc) {
    const float one = 1.0f, zero = 0.0f;
    BH_BLAS(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero,
                          C, ldc));
}

// Wl, Wu (n x n, ld n) = the inverses of A's unit-lower and non-unit upper
// triangles, each with its other triangle zero, ready for a GEMM (either may
// be null; both are built by the same launches)
template <class T>
static void tri_inverse(Solver* ctx, Lane L, int n, const T* A, int lda, T* Wl, T* Wu) {
    T* t = ctx->lu.tinv.as<T>();
    tri_inv2<T>(L.s, n, A, lda, Wl, Wu, t, t + tri_inv_tmp_elems(n));
}

// max |x| over a column-major block, merged into *out as the bits of a
// non-negative double (NaN orders above every number, so it fails a <= test).
// grid.x: 256-row tiles, grid.y: 64-column chunks.
template <class T>
__global__ __launch_bounds__(256) void k_absmax_bits(int rows, int cols, const T* __restrict__ X, int ld,
                                                     unsigned long long* __restrict__ out) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * 64, c1 = min(cols, c0 + 64);
    unsigned long long bits = 0;
    if (r < rows)
        for (int c = c0; c < c1; ++c) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(fabs((double)X[(size_t)c * ld + r]));
            bits = b > bits ? b : bits;
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, bits);
}

// ---------------------------------------------------- split-fp16 update --
// max |x| over a column-major block (rows x cols, ld) -> bits of a float in *out.
// grid.x: 256-row tiles, grid.y: 64-column chunks.
__global__ __launch_bounds__(256) void k_block_absmax(int rows, int cols, const float* __restrict__ X, int ld,
                                                      unsigned int* __restrict__ out) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * 64, c1 = min(cols, c0 + 64);
    float m = 0.0f;
    if (r < rows)
        for (int c = c0; c < c1; ++c) m = fmaxf(m, fabsf(X[(size_t)c * ld + r]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));  // non-negative floats order as uints
}

// scale exponents: A and B scaled so that max |.| maps to [2^11, 2^12)
__device__ inline int scale_exp(unsigned int bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f)) return 0;
    int e;
    frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return 12 - e;
}
// |L21| <= 1 for every panel that reaches a trailing update (partial pivoting,
// or a speculative panel accepted because max |l| <= 1): its scale is fixed
constexpr int SPLIT_EA = 11;

// Acat (rows x 3K, ld rows) = [hi | hi | lo] of 2^11 A.  With maxl set, max |a|
// is merged into it as the bits of a non-negative double (NaN above every
// number), the check that accepts a speculative panel.
__global__ __launch_bounds__(256) void k_split_a(int rows, int K, const float* __restrict__ A, int lda,
                                                 __half* __restrict__ out, unsigned long long* __restrict__ maxl) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * 64, c1 = min(K, c0 + 64);
    unsigned int mb = 0;   // bits of max |a|: non-negative floats (and NaN above them) order as uints
    if (r < rows)
        for (int c = c0; c < c1; ++c) {
            const float a = A[(size_t)c * lda + r];
            const unsigned int ab = __float_as_uint(fabsf(a));
            mb = ab > mb ? ab : mb;
            const float v = ldexpf(a, SPLIT_EA);
            const __half h = __float2half_rn(v);
            const __half l = __float2half_rn(v - __half2float(h));
            out[(size_t)c * rows + r] = h;
            out[(size_t)(c + K) * rows + r] = h;
            out[(size_t)(c + 2 * K) * rows + r] = l;
        }
    if (maxl == nullptr) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int o = __shfl_xor(mb, off);
        mb = o > mb ? o : mb;
    }
    if ((threadIdx.x & 63) == 0) {
        const unsigned long long bits = mb > 0x7F800000u ? 0x7FF8000000000000ull
                                                         : (unsigned long long)__double_as_longlong((double)__uint_as_float(mb));
        atomicMax(maxl, bits);
    }
}

// Bcat (3K x cols, ld 3K) = [hi; lo; hi] of 2^eb B (K rows), eb from max |B|
// in *mx; work-group (0, 0) also writes the GEMM's alpha = -2^-(11 + eb) and
// beta = 1 to ab[0..1].
__global__ __launch_bounds__(256) void k_split_b(int K, int cols, const float* __restrict__ B, int ldb,
                                                 const unsigned int* __restrict__ mx, __half* __restrict__ out,
                                                 float* __restrict__ ab) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int eb = scale_exp(*mx);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ab[0] = -ldexpf(1.0f, -(SPLIT_EA + eb));
        ab[1] = 1.0f;
    }
    if (r >= K) return;
    const size_t K3 = 3 * (size_t)K;
    const int c0 = blockIdx.y * 64, c1 = min(cols, c0 + 64);
    for (int c = c0; c < c1; ++c) {
        const float v = ldexpf(B[(size_t)c * ldb + r], eb);
        const __half h = __float2half_rn(v);
        const __half l = __float2half_rn(v - __half2float(h));
        out[(size_t)c * K3 + r] = h;
        out[(size_t)c * K3 + K + r] = l;
        out[(size_t)c * K3 + 2 * K + r] = h;
    }
}

// Bcat as k_split_b, with the factorisation's fixed exponent eb; the block's
// max |B| (float bits) goes to *log for the check after the factorisation
__global__ __launch_bounds__(256) void k_split_b_fixed(int K, int cols, const float* __restrict__ B, int ldb, int eb,
                                                       __half* __restrict__ out, unsigned int* __restrict__ log) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (r < K) {
        const size_t K3 = 3 * (size_t)K;
        const int c0 = blockIdx.y * 64, c1 = min(cols, c0 + 64);
        for (int c = c0; c < c1; ++c) {
            const float b = B[(size_t)c * ldb + r];
            m = fmaxf(m, fabsf(b));
            const float v = ldexpf(b, eb);
            const __half h = __float2half_rn(v);
            const __half l = __float2half_rn(v - __half2float(h));
            out[(size_t)c * K3 + r] = h;
            out[(size_t)c * K3 + K + r] = l;
            out[(size_t)c * K3 + 2 * K + r] = h;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(log, __float_as_uint(m));   // NaN bits order above every number
}

// max |A(i, i)| over the diagonal (float bits)
__global__ __launch_bounds__(256) void k_diag_absmax(int N, const float* __restrict__ A, unsigned int* __restrict__ out) {
    float m = 0.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
        m = fmaxf(m, fabsf(A[(size_t)i * N + i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// k_split_a with two rows per thread (rows and lda even): 4-byte stores of
// half pairs instead of 2-byte stores; same values
__global__ __launch_bounds__(256) void k_split_a2(int rows, int K, const float* __restrict__ A, int lda,
                                                  __half* __restrict__ out, unsigned long long* __restrict__ maxl) {
    const int r = 2 * (blockIdx.x * 256 + threadIdx.x);
    const int c0 = blockIdx.y * 32, c1 = min(K, c0 + 32);
    unsigned int mb = 0;
    if (r < rows)
        for (int c = c0; c < c1; ++c) {
            const float2 a = *reinterpret_cast<const float2*>(A + (size_t)c * lda + r);
            const float av[2] = {a.x, a.y};
            unsigned int hb = 0, lb = 0;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const unsigned int ab = __float_as_uint(fabsf(av[u]));
                mb = ab > mb ? ab : mb;
                const float v = ldexpf(av[u], SPLIT_EA);
                const __half h = __float2half_rn(v);
                const __half l = __float2half_rn(v - __half2float(h));
                hb |= (unsigned int)__half_as_ushort(h) << (16 * u);
                lb |= (unsigned int)__half_as_ushort(l) << (16 * u);
            }
            *reinterpret_cast<unsigned int*>(out + (size_t)c * rows + r) = hb;
            *reinterpret_cast<unsigned int*>(out + (size_t)(c + K) * rows + r) = hb;
            *reinterpret_cast<unsigned int*>(out + (size_t)(c + 2 * K) * rows + r) = lb;
        }
    if (maxl == nullptr) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int o = __shfl_xor(mb, off);
        mb = o > mb ? o : mb;
    }
    if ((threadIdx.x & 63) == 0) {
        const unsigned long long bits = mb > 0x7F800000u ? 0x7FF8000000000000ull
                                                         : (unsigned long long)__double_as_longlong((double)__uint_as_float(mb));
        atomicMax(maxl, bits);
    }
}

// k_split_b_fixed with four rows per thread (K and ldb multiples of 4): 16-byte
// loads and 8-byte stores of half quadruples; same values
// When dst is given, B's values also go to dst (ld ldd): the U12 block into
// the factor, in the same pass.
__global__ __launch_bounds__(256) void k_split_b_fixed4(int K, int cols, const float* __restrict__ B, int ldb, int eb,
                                                        __half* __restrict__ out, unsigned int* __restrict__ log,
                                                        float* __restrict__ dst, int ldd) {
    const int r = 4 * (blockIdx.x * 256 + threadIdx.x);
    float m = 0.0f;
    if (r < K) {
        const size_t K3 = 3 * (size_t)K;
        const int c0 = blockIdx.y * 16, c1 = min(cols, c0 + 16);
        for (int c = c0; c < c1; ++c) {
            const float4 b = *reinterpret_cast<const float4*>(B + (size_t)c * ldb + r);
            if (dst) {
                float* d = dst + (size_t)c * ldd + r;   // any ldd: four 4-byte stores
                d[0] = b.x; d[1] = b.y; d[2] = b.z; d[3] = b.w;
            }
            const float bv[4] = {b.x, b.y, b.z, b.w};
            unsigned int h2[2] = {0u, 0u}, l2[2] = {0u, 0u};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                m = fmaxf(m, fabsf(bv[u]));
                const float v = ldexpf(bv[u], eb);
                const __half h = __float2half_rn(v);
                const __half l = __float2half_rn(v - __half2float(h));
                h2[u >> 1] |= (unsigned int)__half_as_ushort(h) << (16 * (u & 1));
                l2[u >> 1] |= (unsigned int)__half_as_ushort(l) << (16 * (u & 1));
            }
            const uint2 hv = make_uint2(h2[0], h2[1]), lv = make_uint2(l2[0], l2[1]);
            *reinterpret_cast<uint2*>(out + (size_t)c * K3 + r) = hv;
            *reinterpret_cast<uint2*>(out + (size_t)c * K3 + K + r) = lv;
            *reinterpret_cast<uint2*>(out + (size_t)c * K3 + 2 * K + r) = hv;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(log, __float_as_uint(m));
}

// fp16 splits of a panel's L21 (rows x k) into the step's A buffer (parity p)
static void split_a(Solver* ctx, hipStream_t st, int rows, int k, const float* A, int lda, int p,
                    unsigned long long* maxl) {
    if (rows <= 0) return;
    __half* acat = ctx->split_a[p].as<__half>();
    if (rows % 2 == 0 && lda % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 8 == 0)
        k_split_a2<<<dim3((rows / 2 + 255) / 256, (k + 31) / 32), 256, 0, st>>>(rows, k, A, lda, acat, maxl);
    else
        k_split_a<<<dim3((rows + 255) / 256, (k + 63) / 64), 256, 0, st>>>(rows, k, A, lda, acat, maxl);
    BH_HIP(hipGetLastError());
}

// fp16 splits of U12 columns [col0, col0 + n) (k x n, ld ldb) with their own
// scale; `slot` (0: the next panel's columns, 1: the rest) selects the scale
// word and the alpha pair, whose scale word the caller has zeroed.
// With dst (ld ldd) the FP32 block B is also copied there when the vectorised
// kernel runs; returns whether it was.
static bool split_b(Solver* ctx, hipStream_t st, int k, int n, const float* B, int ldb, int col0, int slot,
                    float* dst = nullptr, int ldd = 0) {
    if (n <= 0) return true;
    __half* bcat = ctx->split_b.as<__half>() + (size_t)col0 * 3 * k;
    if (ctx->split_fixed_on) {
        BH_CHECK(ctx->split_calls < ctx->split_nlog, BEMHIP_ERR_HIP, "split log overflow");
        unsigned int* log = ctx->split_log.as<unsigned int>() + ctx->split_calls++;
        if (k % 4 == 0 && ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0 &&
            reinterpret_cast<uintptr_t>(bcat) % 8 == 0) {
            k_split_b_fixed4<<<dim3((k / 4 + 255) / 256, (n + 15) / 16), 256, 0, st>>>(
                k, n, B, ldb, ctx->split_eb, bcat, log, dst, ldd);
            BH_HIP(hipGetLastError());
            return dst != nullptr;
        }
        k_split_b_fixed<<<dim3((k + 255) / 256, (n + 63) / 64), 256, 0, st>>>(k, n, B, ldb, ctx->split_eb, bcat,
                                                                             log);
        BH_HIP(hipGetLastError());
        return false;
    }
    unsigned int* mx = ctx->split_mx.as<unsigned int>();
    float* ab = reinterpret_cast<float*>(mx + 4) + 2 * slot;
    k_block_absmax<<<dim3((k + 255) / 256, (n + 63) / 64), 256, 0, st>>>(k, n, B, ldb, mx + slot);
    k_split_b<<<dim3((k + 255) / 256, (n + 63) / 64), 256, 0, st>>>(k, n, B, ldb, mx + slot, bcat, ab);
    BH_HIP(hipGetLastError());
    return false;
}

// C[:, col0:col0+n] -= L21 U12[:, col0:col0+n] from the prepared splits
static void split_gemm(Solver* ctx, rocblas_handle h, int m, int n, int k, int col0, float* C, int ldc, int p,
                       int slot) {
    const __half* acat = ctx->split_a[p].as<__half>();
    const __half* bcat = ctx->split_b.as<__half>() + (size_t)col0 * 3 * k;
    if (ctx->split_fixed_on) {
        // alpha known on the host: no device-pointer scalars (which make
        // rocBLAS copy them back and wait for the stream)
        const float alpha = -std::ldexp(1.0f, -(SPLIT_EA + ctx->split_eb)), beta = 1.0f;
        BH_BLAS(rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_none, m, n, 3 * k, &alpha, acat,
                                rocblas_datatype_f16_r, m, bcat, rocblas_datatype_f16_r, 3 * k, &beta, C,
                                rocblas_datatype_f32_r, ldc, C, rocblas_datatype_f32_r, ldc, rocblas_datatype_f32_r,
                                rocblas_gemm_algo_standard, 0, 0));
        return;
    }
    float* ab = reinterpret_cast<float*>(ctx->split_mx.as<unsigned int>() + 4) + 2 * slot;
    BH_BLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_device));
    rocblas_status s = rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_none, m, n, 3 * k, ab,
                                       acat, rocblas_datatype_f16_r, m, bcat, rocblas_datatype_f16_r, 3 * k, ab + 1,
                                       C, rocblas_datatype_f32_r, ldc, C, rocblas_datatype_f32_r, ldc,
                                       rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
    BH_BLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
    BH_BLAS(s);
}

// timed trailing update C -= A B; with split16 the fp16 splits of A (buffer
// parity p) and of B's columns col0.. (scale slot `slot`) were prepared by
// split_a / split_b and A, B are not read here.
template <class T>
static void trailing_update(Solver* ctx, hipStream_t st, rocblas_handle h, int m, int n, int k, const T* A, int lda,
                            const T* B, int ldb, T* C, int ldc, bool split16, int col0, int p, int slot,
                            const Mm16Tiles* tiles = nullptr) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        // event pairs are recorded now and read once after the factorisation
        const size_t i = ctx->gemm_ev_used;
        while (ctx->gemm_ev.size() < 2 * (i + 1)) {
            hipEvent_t e;
            BH_HIP(hipEventCreate(&e));
            ctx->gemm_ev.push_back(e);
        }
        e0 = ctx->gemm_ev[2 * i];
        e1 = ctx->gemm_ev[2 * i + 1];
        ctx->gemm_fl.resize(i + 1);
        ctx->gemm_fl[i] = 2.0 * (double)m * (double)n * (double)k;
        ctx->gemm_ev_used = i + 1;
        BH_HIP(hipEventRecord(e0, st));
    }
    if constexpr (std::is_same<T, float>::value) {
        if (ctx->lu16)
            mm16_gemm(st, MM_SUB, 0, ctx->lu.t_l21[p].p, k, ctx->lu.t_u12[0].p, k, col0, m, n, k, C, ldc,
                      std::ldexp(1.0f, ctx->lu16_e - 22),
                      ctx->lookahead && ctx->mm_reserve > 0 ? std::max(8, ctx->ncu - ctx->mm_reserve) : 0, tiles);
        else if (split16) split_gemm(ctx, h, m, n, k, col0, C, ldc, p, slot);
        else gemm_minus(h, m, n, k, A, lda, B, ldb, C, ldc);
    } else {
        gemm_minus(h, m, n, k, A, lda, B, ldb, C, ldc);
    }
    if (ctx->timing) BH_HIP(hipEventRecord(e1, st));
}

// sums the recorded trailing-update timings (after the stream has drained)
static void collect_gemm_timing(Solver* ctx) {
    for (size_t i = 0; i < ctx->gemm_ev_used; ++i) {
        float ms = 0.0f;
        BH_HIP(hipEventElapsedTime(&ms, ctx->gemm_ev[2 * i], ctx->gemm_ev[2 * i + 1]));
        ctx->gemm_ms += ms;
        ctx->gemm_flops += ctx->gemm_fl[i];
        ctx->gemm_launches += 1;
    }
    ctx->gemm_ev_used = 0;
}


// ------------------------------------------------------------- fp16 LU --
// Every product of the fp16 LU runs on fp16 matrix cores with FP32
// accumulation (mm16.hip).  Operands are scaled by powers of two (exact) from
// e, max |diag A| in [2^(e-1), 2^e): blocks of the matrix (A21, A12, U12) by
// 2^(8-e), U11^-1 by 2^(e-2), L11^-1 by 1 and L21 (|l| <= 1) by 2^14; every
// pack logs its largest scaled magnitude, and a factorisation whose log
// exceeds 2^15 (or is not finite) is reported as an fp16 overflow.
static float lu16_sa(const Solver* ctx) { return std::ldexp(1.0f, 8 - ctx->lu16_e); }

// Step p's products on the panel stream: L21 = A21 U11^-1 (rows below an
// m x n panel, A21 at A + n) into l21 (ld m - n) with its A tiles and max |l|
// merged into maxl (when set); with l21 == A + n (a panel factored with row
// interchanges, L21 already in place) only its tiles, and L11^-1 (which the
// interchange-free panel had built with U11^-1).  Then L11^-1 as A tiles.
static void lu16_panel_products(Solver* ctx, Lane L, int m, int n, const float* A, int lda, float* l21, int ldl,
                                int p, unsigned long long* maxl) {
    unsigned int* log = ctx->lu.f16log.as<unsigned int>();
    if (l21 != A + n) {
        mm16_pack(L.s, true, n, n, ctx->lu.winv_u.as<float>(), n, std::ldexp(1.0f, ctx->lu16_e - 2), ctx->lu.t_wu.p,
                  log);
        // A21's tiles, unless the update that produced A21 wrote them
        if (!ctx->lu16_a21_ready) mm16_pack(L.s, false, m - n, n, A + n, lda, lu16_sa(ctx), ctx->lu.t_a21.p, log);
        const Mm16Tiles tl{ctx->lu.t_l21[p].p, false, n, 0, 16384.0f, log, maxl};
        mm16_gemm(L.s, MM_STORE, 2, ctx->lu.t_a21.p, n, ctx->lu.t_wu.p, n, 0, m - n, n, n, l21, ldl,
                  std::ldexp(1.0f, -6), 0, &tl);
    } else {
        mm16_pack(L.s, false, m - n, n, l21, ldl, 16384.0f, ctx->lu.t_l21[p].p, log, nullptr, 0, maxl);
        tri_inverse<float>(ctx, L, n, A, lda, ctx->lu.winv_l[p].as<float>(), nullptr);
    }
    mm16_pack(L.s, false, n, n, ctx->lu.winv_l[p].as<float>(), n, 1.0f, ctx->lu.t_wl[p].p, log);
}

// columns [c0, c0 + nc) of step p's block row (jb rows at A12, ld lda):
// U12 = L11^-1 A12 into the factor with its B tiles, then the trailing update
// of those columns (rest rows below the block row).  a12_tiled: A12's tiles
// were written by the previous step's update; upd: tiles the update writes of
// its result for the next step (the next panel's A21 or the next block row).
static void lu16_cols(Solver* ctx, Lane L, int jb, int rest, int c0, int nc, float* A12, int lda, int p,
                      bool a12_tiled = false, const Mm16Tiles* upd = nullptr) {
    if (nc <= 0) return;
    unsigned int* log = ctx->lu.f16log.as<unsigned int>();
    const size_t off = (size_t)(c0 / 32) * (mm_pad(jb, MM_BK) / 16) * 1024;   // bytes of c0 / 32 column blocks
    float* a12 = A12 + (size_t)c0 * lda;
    if (!a12_tiled) mm16_pack(L.s, true, nc, jb, a12, lda, lu16_sa(ctx), ctx->lu.t_a12.as<char>() + off, log);
    const Mm16Tiles tu{ctx->lu.t_u12[0].p, true, jb, c0, lu16_sa(ctx), log, nullptr};
    mm16_gemm(L.s, MM_STORE, 1, ctx->lu.t_wl[p].p, jb, ctx->lu.t_a12.p, jb, c0, jb, nc, jb, a12, lda,
              std::ldexp(1.0f, ctx->lu16_e - 8), 0, &tu);
    trailing_update<float>(ctx, L.s, L.h, rest, nc, jb, nullptr, 0, nullptr, 0, a12 + jb, lda, false, c0, p, 0, upd);
}

// ------------------------------------------------------------ factorise --
template <class T>
static void leaf(Solver* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base) {
    const int P = (m + PROWS - 1) / PROWS;
    BH_CHECK(P <= MAXP, BEMHIP_ERR_ARG, "matrix too large for the panel kernel (N > 65536)");
    char* ws = ctx->lu.panel_ws.as<char>();
    // P <= 256 work-groups of 256 threads always fit on the device at once (1-2 per
    // CU by LDS), so a plain launch is co-resident once any concurrent kernel on
    // the other stream retires work-groups; the barrier's bounded spin reports
    // (never hangs) if that assumption were broken.
    k_panel<T><<<P, PROWS, 0, L.s>>>(m, n, A, lda, ipiv, roff, ws, base);
    BH_HIP(hipGetLastError());
    base += (unsigned)n * (unsigned)P;
}

template <class T>
static void laswp(Lane L, int ncols, T* A, int lda, const int* ipiv, int k1, int k2, int roff) {
    if (ncols <= 0 || k2 <= k1) return;
    k_laswp<T><<<(ncols + 255) / 256, 256, 0, L.s>>>(ncols, A, lda, ipiv, k1, k2, roff);
    BH_HIP(hipGetLastError());
}

template <class T>
static void panel_rec(Solver* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base) {
    if (n <= LEAF) { leaf(ctx, L, m, n, A, lda, ipiv, roff, base); return; }
    int n1 = ((n / 2 + LEAF - 1) / LEAF) * LEAF;
    if (n1 >= n) n1 = n - LEAF;
    const int n2 = n - n1;
    panel_rec(ctx, L, m, n1, A, lda, ipiv, roff, base);
    T* A12 = A + (size_t)n1 * lda;
    laswp(L, n2, A12, lda, ipiv, 0, n1, roff);
    trsm_llnu(L.h, n1, n2, A, lda, A12, lda);
    if (m > n1) {
        gemm_minus(L.h, m - n1, n2, n1, A + n1, lda, A12, lda, A12 + n1, lda);
        panel_rec(ctx, L, m - n1, n2, A12 + n1, lda, ipiv + n1, roff + n1, base);
        laswp(L, n1, A + n1, lda, ipiv + n1, 0, n2, roff + n1);
    }
}

// Interchange-free factorisation of an m x n panel: the top n x n block by
// 32-column steps (k_np_step, then the rank-32 update of the rest of the top
// block, k_np_update), then every row below it at once, L21 = A21 U11^-1 (the
// explicit inverse of U11 in one GEMM) written to `l21` ((m - n) x n, ld
// m - n), not into A: A21 keeps its values until the panel is accepted, so a
// rejected panel only has to restore its top block.  Without interchanges
// this is the same elimination as column-by-column Gaussian elimination of the
// whole panel.  (A single grid-barrier launch for the top block was measured
// slower: its spinning work-groups take compute units from the concurrent
// trailing-update GEMM.)
template <class T>
static void np_steps(Lane L, int n, T* A, int lda, unsigned long long* maxl, T* dscr);

// the FP32 top block in one launch with the stripes in registers (it also
// writes the backup of the block)
template <class T> static bool top32(const Solver* ctx, int n) {
    return std::is_same<T, float>::value && ctx->np_top && ctx->np_top32 && (n + NPB - 1) / NPB <= TOP_STRIPES &&
           n <= top_max_rows<T>();
}

template <class T>
static void panel_np(Solver* ctx, Lane L, int m, int n, T* A, int lda, unsigned int* sync, T* l21, int p) {
    unsigned long long* maxl = reinterpret_cast<unsigned long long*>(sync);
    const int stripes = (n + NPB - 1) / NPB;
    if (top32<T>(ctx, n)) {
        if constexpr (std::is_same<T, float>::value)
            k_np_top32<<<stripes, top_threads<T>(n), 0, L.s>>>(n, A, lda, sync, ++ctx->np_epoch, ctx->lu.diag.as<T>(),
                                                               ctx->lu.backup.as<T>(), nullptr, stripes, 0, 0, 0,
                                                               nullptr);
    } else if (ctx->np_top && stripes <= TOP_STRIPES && n <= top_max_rows<T>()) {
        T* linv = ctx->lu.diag.as<T>();
        k_np_top<T><<<stripes, top_threads<T>(n), 0, L.s>>>(n, A, lda, sync, ++ctx->np_epoch, linv);
    } else {
        T* dscr = ctx->lu.diag.as<T>();
        np_steps(L, n, A, lda, maxl, dscr);
    }
    BH_HIP(hipGetLastError());
    if (m > n) {
        // U11^-1 for L21 and L11^-1 for the step's U12, in the same launches
        T* W = ctx->lu.winv_u.as<T>();
        tri_inverse<T>(ctx, L, n, A, lda, ctx->lu.winv_l[p].as<T>(), W);
        if (ctx->lu16) return;   // L21 with the step's other products (lu16_panel_products)
        gemm_nn(L.h, m - n, n, n, A + n, lda, W, n, l21, m - n);
    }
}

// the top block by 32-column steps: k_np_step, then the rank-32 update of the
// rest of the top block (k_np_update)
template <class T>
static void np_steps(Lane L, int n, T* A, int lda, unsigned long long* maxl, T* dscr) {
    for (int j0 = 0; j0 < n; j0 += NPB) {
        const int bs = std::min(NPB, n - j0);
        const int rows = n - j0, ncols = n - j0 - bs;
        T* Ajj = A + (size_t)j0 * lda + j0;
        const int prow = (rows - bs + 255) / 256, pcol = (ncols + 255) / 256;
        k_np_step<T><<<std::max(1, prow + pcol), 256, 0, L.s>>>(rows, bs, ncols, prow, Ajj, lda, maxl, dscr);
        if (rows > bs && ncols > 0)
            k_np_update<T><<<dim3((rows - bs + 63) / 64, (ncols + 63) / 64), 256, 0, L.s>>>(
                rows - bs, ncols, bs, Ajj + bs, Ajj + (size_t)bs * lda, Ajj + (size_t)bs * lda + bs, lda, dscr, Ajj);
        else
            k_np_diag_store<T><<<1, 256, 0, L.s>>>(bs, dscr, Ajj, lda);
    }
}

// What a panel's lane leaves for the main stream's update of its step: the
// inverse of the panel's unit-lower diagonal block (U12 = L11^-1 A12 is one
// GEMM with it) and, for the split-fp16 update, the fp16 splits of L21 (whose
// pass also merges max |l| into maxl when that is set).  Both alternate
// between two buffers by step parity p: the side stream makes step k+1's
// while the main stream still reads step k's.
template <class T>
static void panel_products(Solver* ctx, Lane L, int m, int n, const T* A, int lda, const T* l21, int ldl, int p,
                           bool split16, unsigned long long* maxl) {
    if (m <= n) return;   // the last panel: no update follows
    if constexpr (std::is_same<T, float>::value) {
        if (ctx->lu16) {
            lu16_panel_products(ctx, L, m, n, A, lda, const_cast<float*>(l21), ldl, p, maxl);
            return;
        }
    }
    // L11^-1 of a panel factored with interchanges (an interchange-free one
    // had it built with U11^-1)
    if (l21 == A + n) tri_inverse<T>(ctx, L, n, A, lda, ctx->lu.winv_l[p].as<T>(), nullptr);
    if constexpr (std::is_same<T, float>::value) {
        if (split16) {
            split_a(ctx, L.s, m - n, n, l21, ldl, p, maxl);
            return;
        }
    }
    if (maxl) {
        k_absmax_bits<T><<<dim3((m - n + 255) / 256, (n + 63) / 64), 256, 0, L.s>>>(m - n, n, l21, ldl, maxl);
        BH_HIP(hipGetLastError());
    }
}

// Panel of one outer step (m x n at A).  With speculation on, the top block is
// saved, the panel factored without interchanges and the largest multiplier
// copied to the host; the caller enqueues independent work and then calls
// panel_finish, which keeps the result when max |l| <= 1 (identical to GEPP's
// choices) and otherwise restores the top block and factors the panel with
// partial pivoting.
// verdict set: the verdict is not read back; the largest multiplier and the
// error word are merged into verdict (max |l| bits in words 0-1, error in
// word 2) on the lane's stream, for the caller to read once at the end.
__global__ void k_verdict_merge(const unsigned int* __restrict__ sync, unsigned int* __restrict__ verdict) {
    if (threadIdx.x == 0) {
        const unsigned long long ml = (unsigned long long)sync[0] | ((unsigned long long)sync[1] << 32);
        const unsigned long long vm = (unsigned long long)verdict[0] | ((unsigned long long)verdict[1] << 32);
        if (ml > vm) {
            verdict[0] = (unsigned int)ml;
            verdict[1] = (unsigned int)(ml >> 32);
        }
        verdict[2] |= sync[3];
    }
}
template <class T>
static void panel_start(Solver* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base,
                        int p, bool split16, unsigned int* verdict = nullptr) {
    if (!ctx->speculate) {
        panel_rec(ctx, L, m, n, A, lda, ipiv, roff, base);
        panel_products(ctx, L, m, n, A, lda, A + n, lda, p, split16, nullptr);
        return;
    }
    if (!top32<T>(ctx, n)) copy2d(L.s, n, n, A, lda, ctx->lu.backup.as<T>(), n);
    // sync words: max |l| (bits of a non-negative double), ticket, error, flags
    unsigned int* sync = ctx->lu.maxl.as<unsigned int>();
    k_fill_u32<<<1, 64, 0, L.s>>>(sync, 4, 0u);
    T* l21 = ctx->lu.l21[p].as<T>();
    panel_np(ctx, L, m, n, A, lda, sync, l21, p);
    panel_products(ctx, L, m, n, A, lda, l21, m - n, p, split16, reinterpret_cast<unsigned long long*>(sync));
    if (verdict) {
        k_verdict_merge<<<1, 64, 0, L.s>>>(sync, verdict);
        BH_HIP(hipGetLastError());
        return;
    }
    BH_HIP(hipMemcpyAsync(ctx->host_scal + 32, sync, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost, L.s));
    BH_HIP(hipEventRecord(ctx->ev_la[2], L.s));
}

// Returns true when the panel was factored with row interchanges.  *l21 and
// *ldl receive where the step's L21 is: the step's l21 buffer for an accepted
// speculative panel (the caller copies it into A once the panel is handed
// over) or A itself.
template <class T>
static bool panel_finish(Solver* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base,
                         int p, bool split16, const T** l21, int* ldl, bool deferred = false) {
    ctx->panels += 1;
    *l21 = A + n;
    *ldl = lda;
    if (!ctx->speculate) {
        ctx->panels_pivoted += 1;
        return true;
    }
    double ml = 0.0;
    if (!deferred) {
        // (deferred: the panel is taken as accepted; the caller checks the
        // merged verdict at the end and redoes the factorisation otherwise)
        BH_HIP(hipEventSynchronize(ctx->ev_la[2]));
        ml = ctx->host_scal[32];
        BH_CHECK(reinterpret_cast<const unsigned int*>(ctx->host_scal + 33)[1] == 0u, BEMHIP_ERR_HIP,
                 "panel top-block factorisation gave up waiting");
    }
    if (ml <= 1.0) {
        k_iota<<<(n + 255) / 256, 256, 0, L.s>>>(ipiv, n, roff);
        BH_HIP(hipGetLastError());
        *l21 = ctx->lu.l21[p].as<T>();
        *ldl = std::max(1, m - n);
        return false;
    }
    copy2d(L.s, n, n, ctx->lu.backup.as<T>(), n, A, lda);
    panel_rec(ctx, L, m, n, A, lda, ipiv, roff, base);
    panel_products(ctx, L, m, n, A, lda, A + n, lda, p, split16, nullptr);
    ctx->panels_pivoted += 1;
    return true;
}

// The look-ahead loop of lu_factor (after panel 0) with two host threads:
// this one enqueues the main stream's work, a second one drives the panels on
// the side stream.  rocBLAS's gemm_ex with alpha in device memory (the
// split-fp16 update) returns only when the stream has drained up to it, so
// with one host thread the next panel could be enqueued only once the update
// before it had started on the device.  Hand-over per step k: this thread
// records ev_ready after the narrow update and posts panel k+1; the other
// waits for ev_ready on the side stream, factors the panel, records ev_panel
// and reports its pivoting and where its L21 is.
template <class T>
static void lu_two_threads(Solver* ctx, int N, T* A, int* ipiv, bool split16, int LU_NB, unsigned int& base,
                           bool pivoted, const T* l21, int ldl) {
    const Lane main{ctx->stream, ctx->blas}, side{ctx->stream2, ctx->blas2};
    hipEvent_t ev_ready = ctx->ev_la[0], ev_panel = ctx->ev_la[1];
    const int lda = N;
    auto width = [&](int rem) { return std::min(LU_NB, rem); };
    struct Job { int k, jb, p; bool stop; };
    struct Done { bool pivoted; const T* l21; int ldl; };
    std::mutex mu;
    std::condition_variable cv;
    bool have_job = false, have_done = false, failed = false;
    Job job{0, 0, 0, false};
    Done done{false, nullptr, 1};
    std::exception_ptr err;
    int device = 0;
    BH_HIP(hipGetDevice(&device));
    auto fail = [&](std::exception_ptr e) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = e;
            failed = true;
        }
        cv.notify_all();
    };
    std::thread panels([&] {
        try {
            BH_HIP(hipSetDevice(device));
            for (;;) {
                Job jo;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return have_job || failed; });
                    if (failed) return;
                    jo = job;
                    have_job = false;
                }
                if (jo.stop) return;
                const int k1 = jo.k + jo.jb, rest = N - k1, jb2 = width(rest);
                T* Ap = A + (size_t)k1 * lda + k1;
                BH_HIP(hipStreamWaitEvent(side.s, ev_ready, 0));
                panel_start(ctx, side, rest, jb2, Ap, lda, ipiv + k1, k1, base, jo.p, split16);
                Done d{false, nullptr, 1};
                d.pivoted = panel_finish(ctx, side, rest, jb2, Ap, lda, ipiv + k1, k1, base, jo.p, split16, &d.l21,
                                         &d.ldl);
                BH_HIP(hipEventRecord(ev_panel, side.s));
                // an accepted panel's L21 into the factor after the hand-over;
                // the side stream's next panel comes after this copy, and so
                // does every later row swap on the main stream
                if (!d.pivoted) copy2d(side.s, rest - jb2, jb2, d.l21, d.ldl, Ap + jb2, lda);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    done = d;
                    have_done = true;
                }
                cv.notify_all();
            }
        } catch (...) {
            fail(std::current_exception());
        }
    });
    auto post = [&](Job j) {
        {
            std::lock_guard<std::mutex> lk(mu);
            job = j;
            have_job = true;
        }
        cv.notify_all();
    };
    // fp16 LU: the updates of step k also write the tiles step k + 1 packs
    // otherwise: the next panel's A21 (narrow update) and the next block row
    // (wide update; valid when the next panel takes no row interchanges)
    bool a12_tiled = false;
    ctx->lu16_a21_ready = false;
    try {
        int k = 0, jb = width(N), p = 0;
        for (;;) {
            BH_HIP(hipStreamWaitEvent(main.s, ev_panel, 0));   // panel k factored (and the L21 copies before it)
            if (pivoted) laswp(main, k, A + k, lda, ipiv + k, 0, jb, k);
            const int rest = N - k - jb;
            if (rest <= 0) break;
            T* A12 = A + (size_t)(k + jb) * lda + k;
            if (pivoted) laswp(main, rest, A12, lda, ipiv + k, 0, jb, k);
            const int jb2 = width(rest);
            const T* W = ctx->lu.winv_l[p].as<T>();
            T* U = ctx->lu.u12.as<T>();   // U12 = L11^-1 A12, jb x rest, ld jb
            if (split16) {
                k_fill_u32<<<1, 64, 0, main.s>>>(ctx->split_mx.as<unsigned int>(), 2, 0u);
                BH_HIP(hipGetLastError());
            }
            // the next panel's columns first, then hand them over.  The split
            // kernel also writes U12 into the factor (A12 is read only by the
            // U12 product before it); otherwise one copy at the end.
            bool in_factor = false;
            bool f16 = false;
            // the block row's tiles from the previous update only if this
            // step's panel kept its rows
            const bool a12_ok = a12_tiled && !pivoted;
            a12_tiled = false;
            if constexpr (std::is_same<T, float>::value) {
                if (ctx->lu16) {
                    // U12 goes into the factor with its tiles; the narrow update
                    // writes the next panel's A21 tiles (rows below its top block)
                    const bool a21 = rest > jb2;
                    const Mm16Tiles ta{ctx->lu.t_a21.p, false, jb2, 0, lu16_sa(ctx), ctx->lu.f16log.as<unsigned int>(),
                                       nullptr, jb2};
                    lu16_cols(ctx, main, jb, rest, 0, jb2, A12, lda, p, a12_ok, a21 ? &ta : nullptr);
                    ctx->lu16_a21_ready = a21;
                    in_factor = f16 = true;
                }
            }
            if (!f16) {
                gemm_nn(main.h, jb, jb2, jb, W, jb, A12, lda, U, jb);
                if constexpr (std::is_same<T, float>::value)
                    if (split16) in_factor = split_b(ctx, main.s, jb, jb2, U, jb, 0, 0, A12, lda);
                trailing_update<T>(ctx, main.s, main.h, rest, jb2, jb, l21, ldl, U, jb, A12 + jb, lda, split16, 0, p, 0);
            }
            BH_HIP(hipEventRecord(ev_ready, main.s));
            post(Job{k, jb, p ^ 1, false});
            if (rest > jb2) {
                T* Uw = U + (size_t)jb2 * jb;
                T* A12w = A12 + (size_t)jb2 * lda;
                if constexpr (std::is_same<T, float>::value)
                    if (f16) {
                        // the wide update writes the next block row's tiles
                        // (its first jb2 rows), used if the next panel keeps its rows
                        const Mm16Tiles tb{ctx->lu.t_a12.p, true, jb2, 0, lu16_sa(ctx),
                                           ctx->lu.f16log.as<unsigned int>(), nullptr};
                        lu16_cols(ctx, main, jb, rest, jb2, rest - jb2, A12, lda, p, a12_ok, &tb);
                        a12_tiled = true;
                    }
                if (!f16) {
                    gemm_nn(main.h, jb, rest - jb2, jb, W, jb, A12w, lda, Uw, jb);
                    if constexpr (std::is_same<T, float>::value)
                        if (split16)
                            in_factor = split_b(ctx, main.s, jb, rest - jb2, Uw, jb, jb2, 1, A12w, lda) && in_factor;
                    trailing_update<T>(ctx, main.s, main.h, rest, rest - jb2, jb, l21, ldl, Uw, jb, A12w + jb, lda,
                                       split16, jb2, p, 1);
                }
            }
            if (!in_factor) copy2d(main.s, jb, rest, U, jb, A12, lda);   // U12 into the factor
            Done d;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return have_done || failed; });
                if (failed) break;
                d = done;
                have_done = false;
            }
            pivoted = d.pivoted;
            l21 = d.l21;
            ldl = d.ldl;
            k += jb;
            jb = jb2;
            p ^= 1;
        }
    } catch (...) {
        fail(std::current_exception());
    }
    post(Job{0, 0, 0, true});
    panels.join();
    ctx->lu16_a21_ready = false;
    if (err) std::rethrow_exception(err);
}

// a timed block product (as trailing_update times its launches): event pair
// and algorithmic flops recorded, read after the factorisation
template <class F> static void timed_gemm(Solver* ctx, hipStream_t st, double flops, F&& launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        const size_t i = ctx->gemm_ev_used;
        while (ctx->gemm_ev.size() < 2 * (i + 1)) {
            hipEvent_t e;
            BH_HIP(hipEventCreate(&e));
            ctx->gemm_ev.push_back(e);
        }
        e0 = ctx->gemm_ev[2 * i];
        e1 = ctx->gemm_ev[2 * i + 1];
        ctx->gemm_fl.resize(i + 1);
        ctx->gemm_fl[i] = flops;
        ctx->gemm_ev_used = i + 1;
        BH_HIP(hipEventRecord(e0, st));
    }
    launch();
    if (ctx->timing) BH_HIP(hipEventRecord(e1, st));
}

// The fp16 LU with the trailing updates of two panels merged (K = 2 NB): the
// bulk update reads and writes the trailing matrix once per pair of panels.
// Three streams: S factors the panels, X runs the narrow products (U12, the
// updates of the next panels' columns and of a panel's block row), M the bulk
// updates.  For a pair (k, k+1), panel k factored:
//   X: U12_k of panel k+1's columns, their update (K = NB) -> S: panel k+1;
//      once the rows of panels k and k+1 have the previous pair's bulk
//      update, the rest of U12_k and panel k+1's block row updated by panel
//      k; after panel k+1, U12_{k+1} of panel k+2's columns;
//   then, in either order (below), the rest of U12_{k+1} and, once the
//      previous bulk update is done, panel k+2's columns by both panels
//      (K = 2 NB) -> S: panel k+2; last panel k+3's columns by both panels;
//   M: every column beyond panel k+3 by both panels (K = 2 NB), while S
//      factors panels k+2 and k+3.  While at least lu16_bulk_first rows are
//      left, M is the longest chain: U12_{k+1} is finished first, and the
//      bulk update is two launches, the rows of panels k+2 and k+3 first
//      (all that the next pair's U12 and block row wait for), then the rows
//      beyond; later panel k+2's columns come first and the bulk update is
//      one launch.
// Panels rotate over four buffer sets, U12 over four (a bulk update reads two
// panels' products while the next pair's are made).  Only the bulk updates on
// M are timed (the gemm statistics): the narrow products belong to the chain
// from panel to panel.  The panels' verdicts are merged on the device and read
// once at the end; returns -1 when a panel needed row interchanges: the caller
// redoes the factorisation step by step from a fresh copy of the matrix
// (partial pivoting moves rows that a deferred update still has to reach).
static int lu16_pairs(Solver* ctx, int N, float* A, int* ipiv, unsigned int& base) {
    const int NB = ctx->nb16, lda = N;
    const Lane lm{ctx->stream, ctx->blas}, ls{ctx->stream2, ctx->blas2}, lx{ctx->stream3, ctx->blas};
    hipEvent_t ev_ready = ctx->ev_la[0], ev_panel = ctx->ev_la[1], ev_x = ctx->ev_m[0], ev_big = ctx->ev_m[1];
    hipEvent_t ev_b1 = ctx->ev_m[2];
    unsigned int* log = ctx->lu.f16log.as<unsigned int>();
    const float sa = lu16_sa(ctx);                              // tiles of A blocks: 2^(8 - e)
    const float au = std::ldexp(1.0f, ctx->lu16_e - 8);         // U12 product
    const float aup = std::ldexp(1.0f, ctx->lu16_e - 22);       // updates (L21 2^14, U12 2^(8 - e))
    const int maxwg = ctx->mm_reserve > 0 ? std::max(8, ctx->ncu - ctx->mm_reserve) : 0;
    auto width = [&](int rem) { return std::min(NB, rem); };
    // X and M start after everything enqueued on M so far
    BH_HIP(hipEventRecord(ev_big, lm.s));
    BH_HIP(hipStreamWaitEvent(lx.s, ev_big, 0));
    BH_HIP(hipStreamWaitEvent(ls.s, ev_big, 0));
    // A tiles of L21 of panel buffer p (k extent kp) from its row r (a multiple of 32) on
    auto l21_rows = [&](int p, int kp, int r) {
        return static_cast<const void*>(ctx->lu.t_l21[p].as<char>() + (size_t)(r / 32) * (mm_pad(kp, MM_BK) / 16) * 1024);
    };
    // U12 of the block row (jb rows at A12, ld lda) for columns [c0, c0 + nc),
    // into the factor and as B tiles into t_u12[u] (from column c0), on X
    auto u12 = [&](int jb, float* A12, int c0, int nc, int p, int u, bool tiled) {
        if (nc <= 0) return;
        float* a12 = A12 + (size_t)c0 * lda;
        if (!tiled) {
            const size_t off = (size_t)(c0 / 32) * (mm_pad(jb, MM_BK) / 16) * 1024;
            mm16_pack(lx.s, true, nc, jb, a12, lda, sa, ctx->lu.t_a12.as<char>() + off, log);
        }
        const Mm16Tiles tu{ctx->lu.t_u12[u].p, true, jb, c0, sa, log, nullptr};
        mm16_gemm(lx.s, MM_STORE, 1, ctx->lu.t_wl[p].p, jb, ctx->lu.t_a12.p, jb, c0, jb, nc, jb, a12, lda, au, 0,
                  &tu);
    };
    // a factored panel: the verdict (host), its L21 into the factor on S, and
    // X continues after it
    const float* l21 = nullptr;
    int ldl = 1;
    auto finish = [&](int m, int n, float* Ap, int* ip, int roff, int p) {
        if (panel_finish(ctx, ls, m, n, Ap, lda, ip, roff, base, p, false, &l21, &ldl, true)) return false;
        BH_HIP(hipEventRecord(ev_panel, ls.s));
        BH_HIP(hipStreamWaitEvent(lx.s, ev_panel, 0));
        copy2d(ls.s, m - n, n, l21, ldl, Ap + n, lda);
        return true;
    };
    int k = 0, jb = width(N), pk = 0;
    ctx->lu16_a21_ready = false;
    // the panels' verdicts are merged on the device and read once at the end
    // (no host round trip per panel); a panel that needed interchanges makes
    // the whole factorisation a redo either way
    unsigned int* verdict = static_cast<unsigned int*>(ctx->lu.verdict.ensure(16));
    k_fill_u32<<<1, 64, 0, ls.s>>>(verdict, 4, 0u);
    panel_start(ctx, ls, N, jb, A, lda, ipiv, 0, base, pk, false, verdict);
    if (!finish(N, jb, A, ipiv, 0, pk)) return -1;
    bool a12_tiled = false;
    // the previous pair's bulk update is on M in two launches: the rows of the
    // next two panels (ev_b1 marks its end), then the rows beyond (ev_big)
    bool big = false;
    for (;;) {
        const int rest = N - k - jb;
        if (rest <= 0) break;
        const int jb2 = width(rest), rest2 = rest - jb2, p1 = (pk + 1) % 4;
        float* A12 = A + (size_t)(k + jb) * lda + k;   // panel k's block row (jb rows)
        float* C1 = A12 + jb;                           // rows and columns beyond panel k
        // panel k+1's columns (updated by the previous pair on X already)
        u12(jb, A12, 0, jb2, pk, pk, a12_tiled);
        {
            const Mm16Tiles ta{ctx->lu.t_a21.p, false, jb2, 0, sa, log, nullptr, jb2};
            mm16_gemm(lx.s, MM_SUB, 0, ctx->lu.t_l21[pk].p, jb, ctx->lu.t_u12[pk].p, jb, 0, rest, jb2, jb, C1,
                          lda, aup, 0, rest2 > 0 ? &ta : nullptr);
        }
        ctx->lu16_a21_ready = rest2 > 0;
        BH_HIP(hipEventRecord(ev_ready, lx.s));
        BH_HIP(hipStreamWaitEvent(ls.s, ev_ready, 0));
        panel_start(ctx, ls, rest, jb2, C1, lda, ipiv + k + jb, k + jb, base, p1, false, verdict);
        // the rest of the block row and panel k+1's block row: their rows are
        // the first part of the previous pair's bulk update
        if (big) BH_HIP(hipStreamWaitEvent(lx.s, ev_b1, 0));
        u12(jb, A12, jb2, rest2, pk, pk, a12_tiled);
        if (rest2 > 0) {
            const Mm16Tiles tb{ctx->lu.t_a12.p, true, jb2, 0, sa, log, nullptr};
            mm16_gemm(lx.s, MM_SUB, 0, ctx->lu.t_l21[pk].p, jb, ctx->lu.t_u12[pk].p, jb, jb2, jb2, rest2, jb,
                          C1 + (size_t)jb2 * lda, lda, aup, 0, &tb);
        }
        if (!finish(rest, jb2, C1, ipiv + k + jb, k + jb, p1)) return -1;
        if (rest2 <= 0) break;   // panel k+1 was the last
        const int jb3 = width(rest2), rest3 = rest2 - jb3, p2 = (pk + 2) % 4;
        float* A12b = C1 + (size_t)jb2 * lda;   // panel k+1's block row (jb2 rows)
        float* C2 = A12b + jb2;                 // rows and columns beyond panel k+1
        const Mm16Seg sg_n{ctx->lu.t_l21[p1].p, jb2, ctx->lu.t_u12[p1].p, jb2, 0, jb2};
        // every column beyond panel k+3, by both panels, on M: first the rows
        // of panels k+2 and k+3 (with panel k+2's block row tiles of those
        // columns, from column jb4 of t_a12), then the rows beyond; it needs
        // nothing from X beyond U12_{k+1}
        // while the trailing matrix is large the bulk updates are the longest
        // chain: U12_{k+1} is finished first so that M goes on at once; later
        // the panels are, and panel k+2 comes first
        const bool bulk_first = (long long)rest2 >= ctx->lu16_bulk_first;
        auto bulk = [&] {
            BH_HIP(hipEventRecord(ev_x, lx.s));
            const int jb4 = rest3 > 0 ? width(rest3) : 0, rest4 = rest3 - jb4;
            if (rest3 <= 0 || rest4 <= 0) return;
            BH_HIP(hipStreamWaitEvent(lm.s, ev_x, 0));
            // once the panels are the longer chain, nothing waits for the
            // first part alone: one launch over all the rows
            const int r1 = bulk_first ? jb3 + jb4 : rest2;
            const Mm16Tiles tb{ctx->lu.t_a12.p, true, jb3, jb4, sa, log, nullptr};
            const Mm16Seg sg1{ctx->lu.t_l21[p1].p, jb2, ctx->lu.t_u12[p1].p, jb2, jb3 + jb4, jb2};
            timed_gemm(ctx, lm.s, 2.0 * r1 * rest4 * (jb + jb2), [&] {
                mm16_gemm(lm.s, MM_SUB, 0, l21_rows(pk, jb, jb2), jb, ctx->lu.t_u12[pk].p, jb, jb2 + jb3 + jb4, r1,
                          rest4, jb, C2 + (size_t)(jb3 + jb4) * lda, lda, aup, maxwg, &tb, nullptr, &sg1);
            });
            BH_HIP(hipEventRecord(ev_b1, lm.s));
            if (r1 == rest2) {
                BH_HIP(hipEventRecord(ev_big, lm.s));
                big = true;
                return;
            }
            const Mm16Seg sg2{l21_rows(p1, jb2, r1), jb2, ctx->lu.t_u12[p1].p, jb2, jb3 + jb4, jb2};
            timed_gemm(ctx, lm.s, 2.0 * rest4 * rest4 * (jb + jb2), [&] {
                mm16_gemm(lm.s, MM_SUB, 0, l21_rows(pk, jb, jb2 + r1), jb, ctx->lu.t_u12[pk].p, jb, jb2 + jb3 + jb4,
                          rest4, rest4, jb, C2 + r1 + (size_t)(jb3 + jb4) * lda, lda, aup, maxwg, nullptr, nullptr,
                          &sg2);
            });
            BH_HIP(hipEventRecord(ev_big, lm.s));
            big = true;
        };
        // panel k+2's columns by both panels (writing panel k+2's A21 tiles),
        // then panel k+2 on S; its rows beyond panel k+1 are the previous bulk
        // update's second launch
        auto narrow2 = [&] {
            const Mm16Tiles ta{ctx->lu.t_a21.p, false, jb3, 0, sa, log, nullptr, jb3};
            mm16_gemm(lx.s, MM_SUB, 0, l21_rows(pk, jb, jb2), jb, ctx->lu.t_u12[pk].p, jb, jb2, rest2, jb3, jb, C2,
                          lda, aup, 0, rest3 > 0 ? &ta : nullptr, nullptr, &sg_n);
            ctx->lu16_a21_ready = rest3 > 0;
            BH_HIP(hipEventRecord(ev_ready, lx.s));
            BH_HIP(hipStreamWaitEvent(ls.s, ev_ready, 0));
            panel_start(ctx, ls, rest2, jb3, C2, lda, ipiv + k + jb + jb2, k + jb + jb2, base, p2, false, verdict);
        };
        u12(jb2, A12b, 0, jb3, p1, p1, true);
        const bool had_big = big;
        big = false;
        if (bulk_first) {
            u12(jb2, A12b, jb3, rest3, p1, p1, true);   // the rest of U12_{k+1}
            if (had_big) BH_HIP(hipStreamWaitEvent(lx.s, ev_big, 0));
            bulk();
            narrow2();
        } else {
            if (had_big) BH_HIP(hipStreamWaitEvent(lx.s, ev_big, 0));
            narrow2();
            u12(jb2, A12b, jb3, rest3, p1, p1, true);
            bulk();
        }
        a12_tiled = false;
        if (rest3 > 0) {
            const int jb4 = width(rest3);
            // panel k+3's columns by both panels (panel k+2's block row tiles
            // of those columns into t_a12 from column 0)
            const Mm16Tiles tb{ctx->lu.t_a12.p, true, jb3, 0, sa, log, nullptr};
            const Mm16Seg sg{ctx->lu.t_l21[p1].p, jb2, ctx->lu.t_u12[p1].p, jb2, jb3, jb2};
            mm16_gemm(lx.s, MM_SUB, 0, l21_rows(pk, jb, jb2), jb, ctx->lu.t_u12[pk].p, jb, jb2 + jb3, rest2, jb4, jb,
                          C2 + (size_t)jb3 * lda, lda, aup, 0, &tb, nullptr, &sg);
            a12_tiled = true;
        }
        if (!finish(rest2, jb3, C2, ipiv + k + jb + jb2, k + jb + jb2, p2)) return -1;
        k += jb + jb2;
        jb = jb3;
        pk = p2;
    }
    // M continues after all three streams
    BH_HIP(hipEventRecord(ev_x, lx.s));
    BH_HIP(hipStreamWaitEvent(lm.s, ev_x, 0));
    BH_HIP(hipEventRecord(ev_panel, ls.s));
    BH_HIP(hipStreamWaitEvent(lm.s, ev_panel, 0));
    ctx->lu16_a21_ready = false;
    BH_HIP(hipMemcpyAsync(ctx->host_scal + 34, verdict, 16, hipMemcpyDeviceToHost, ls.s));
    BH_HIP(hipStreamSynchronize(ls.s));
    BH_CHECK(reinterpret_cast<const unsigned int*>(ctx->host_scal + 35)[0] == 0u, BEMHIP_ERR_HIP,
             "panel top-block factorisation gave up waiting");
    const double ml = ctx->host_scal[34];
    if (!(ml <= 1.0)) return -1;
    return 0;
}

// Blocked right-looking LU with partial pivoting and one block of look-ahead.
// Per outer step k (panel factored, its L21 and L11^-1 ready): the main stream
// forms U12 = L11^-1 A12 and updates the next panel's columns first, hands them
// to the second stream, which factors panel k+1 (and prepares its L11^-1 and
// L21 splits) while the main stream forms the rest of U12 and applies the
// rest of the update.  Only the narrow part of step k lies on the chain of
// dependent work from panel to panel.
template <class T> int lu_factor(Solver* ctx, int N, T* A, int* ipiv, bool split16) {
    // outer block: the split-fp16 update runs at fp16 MFMA speed, so its K must
    // be larger for the C read/write of each update to stay below the compute
    const bool f16 = std::is_same<T, float>::value && ctx->lu16;
    const int LU_NB = f16 ? ctx->nb16 : split16 ? ctx->nb_split : (sizeof(T) == 8 ? ctx->nb64 : ctx->nb32);
    BH_CHECK(!f16 || LU_NB % MM_TILE == 0, BEMHIP_ERR_ARG, "fp16 LU block must be a multiple of 256");
    BH_CHECK(LU_NB >= LEAF && LU_NB % LEAF == 0, BEMHIP_ERR_ARG, "LU block must be a multiple of 32");
    ctx->lu16_a21_ready = false;   // the first panel packs its own A21
    ctx->lu.panel_ws.ensure(PanelWs::BYTES);
    char* ws = ctx->lu.panel_ws.as<char>();
    // work buffers at their largest (the first step's) before anything is
    // enqueued: growing one later would free memory under running kernels
    const size_t nb = (size_t)std::min(LU_NB, N);
    for (int p = 0; p < 4; ++p) {
        ctx->lu.winv_l[p].ensure(sizeof(T) * nb * nb);
        ctx->lu.l21[p].ensure(sizeof(T) * (size_t)N * nb);
        if (split16) ctx->split_a[p].ensure(sizeof(__half) * 3 * (size_t)N * nb);
    }
    ctx->lu.winv_u.ensure(sizeof(T) * nb * nb);
    ctx->lu.tinv.ensure(sizeof(T) * 2 * tri_inv_tmp_elems((int)nb));
    ctx->lu.u12.ensure(sizeof(T) * nb * (size_t)N);
    ctx->lu.backup.ensure(sizeof(T) * nb * nb);
    // published diagonal-block inverses (persistent top-block launches) or the
    // step kernels' diagonal block, at the largest size for either precision
    ctx->lu.diag.ensure(sizeof(double) * NPB * NPB * TOP_STRIPES);
    if (ctx->lu.maxl.bytes < sizeof(unsigned int) * TOP_SYNC_WORDS) {
        // flags start below every publication stamp
        ctx->lu.maxl.ensure(sizeof(unsigned int) * TOP_SYNC_WORDS);
        k_fill_u32<<<1, 128, 0, ctx->stream>>>(ctx->lu.maxl.as<unsigned int>(), TOP_SYNC_WORDS, 0u);
        BH_HIP(hipGetLastError());
        ctx->np_epoch = 0;
    }
    if (split16) {
        ctx->split_b.ensure(sizeof(__half) * 3 * nb * (size_t)N);
        ctx->split_mx.ensure(64);
    }
    if (f16) {
        ctx->lu.t_a21.ensure(mm_tile_bytes(N, (int)nb));
        ctx->lu.t_wu.ensure(mm_tile_bytes((int)nb, (int)nb));
        for (int p = 0; p < 4; ++p) {
            ctx->lu.t_wl[p].ensure(mm_tile_bytes((int)nb, (int)nb));
            ctx->lu.t_l21[p].ensure(mm_tile_bytes(N, (int)nb));
        }
        ctx->lu.t_a12.ensure(mm_tile_bytes(N + MM_TILE, (int)nb));
        for (int u = 0; u < 4; ++u) ctx->lu.t_u12[u].ensure(mm_tile_bytes(N + MM_TILE, (int)nb));
        unsigned int* lg = static_cast<unsigned int*>(ctx->lu.f16log.ensure(256));
        // scale exponent from max |diag A|: its max maps to [2^7, 2^8)
        k_fill_u32<<<1, 64, 0, ctx->stream>>>(lg, 2, 0u);
        k_diag_absmax<<<64, 256, 0, ctx->stream>>>(N, reinterpret_cast<const float*>(A), lg + 1);
        BH_HIP(hipGetLastError());
        unsigned int mb = 0;
        BH_HIP(hipMemcpyAsync(&mb, lg + 1, sizeof(mb), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        float dm;
        std::memcpy(&dm, &mb, sizeof(dm));
        int e = 0;
        if (dm > 0.0f && std::isfinite(dm)) std::frexp(dm, &e);
        ctx->lu16_e = std::max(-100, std::min(100, e));
        ctx->lu16_overflow = false;
    }
    ctx->split_fixed_on = false;
    ctx->split_redo = false;
    ctx->split_calls = 0;
    if constexpr (std::is_same<T, float>::value) {
        if (split16 && ctx->split_fixed) {
            // one B scale for the whole factorisation, from max |diag A|: its
            // max maps to [2^9, 2^10); a U12 block is covered while its max
            // scales into [2^-8, 2^15].  The top is the fp16 range with a
            // margin.  Below it the low halves lose bits relative to the
            // block's own max, but their absolute error (2^-24 after scaling,
            // ~2^-34 of max |diag A|) stays far under the FP32 rounding of the
            // entries the update lands on; the floor only catches a matrix
            // whose blocks all shrink by orders of magnitude.
            const int nlog = 2 * (N / LU_NB + 2);
            unsigned int* log = static_cast<unsigned int*>(ctx->split_log.ensure(sizeof(unsigned int) * (nlog + 1)));
            k_fill_u32<<<1, 128, 0, ctx->stream>>>(log, nlog + 1, 0u);
            k_diag_absmax<<<64, 256, 0, ctx->stream>>>(N, A, log + nlog);
            BH_HIP(hipGetLastError());
            unsigned int mb = 0;
            BH_HIP(hipMemcpyAsync(&mb, log + nlog, sizeof(mb), hipMemcpyDeviceToHost, ctx->stream));
            BH_HIP(hipStreamSynchronize(ctx->stream));
            float dm;
            std::memcpy(&dm, &mb, sizeof(dm));
            int e = 0;
            if (dm > 0.0f && std::isfinite(dm)) std::frexp(dm, &e);
            if (dm > 0.0f && std::isfinite(dm) && e > -90 && e < 90) {
                ctx->split_eb = 10 - e;
                ctx->split_nlog = nlog;
                ctx->split_fixed_on = true;
            }
        }
    }
    const Lane main{ctx->stream, ctx->blas};
    const Lane side = ctx->lookahead ? Lane{ctx->stream2, ctx->blas2} : main;
    hipEvent_t ev_ready = ctx->ev_la[0], ev_panel = ctx->ev_la[1], ev_copied = ctx->ev_la[3];
    k_fill_u32<<<1, 64, 0, main.s>>>(reinterpret_cast<unsigned int*>(ws), 16, 0u);
    BH_HIP(hipGetLastError());
    unsigned int base = 0;
    const int lda = N;
    auto width = [&](int rem) { return std::min(LU_NB, rem); };
    bool merged = false;
    if constexpr (std::is_same<T, float>::value) {
        if (f16 && ctx->lu16_merge && ctx->lookahead && ctx->speculate && ctx->la_min_rest == 0 && N > LU_NB) {
            merged = true;
            if (lu16_pairs(ctx, N, A, ipiv, base) < 0) {
                // a panel needed interchanges: the caller starts again step by step
                BH_HIP(hipStreamSynchronize(ctx->stream2));
                BH_HIP(hipStreamSynchronize(main.s));
                collect_gemm_timing(ctx);
                ctx->lu16_a21_ready = false;
                return -2;
            }
        }
    }
    if (!merged) {
    // panel 0: nothing to overlap it with
    const int jb0 = width(N);
    int p = 0;
    const T* l21 = nullptr;
    int ldl = 1;
    panel_start(ctx, main, N, jb0, A, lda, ipiv, 0, base, p, split16);
    bool pivoted = panel_finish(ctx, main, N, jb0, A, lda, ipiv, 0, base, p, split16, &l21, &ldl);
    BH_HIP(hipEventRecord(ev_panel, main.s));
    if (!pivoted) copy2d(main.s, N - jb0, jb0, l21, ldl, A + jb0, lda);
    if (ctx->lookahead && ctx->la_min_rest == 0 && N > jb0) {
        lu_two_threads(ctx, N, A, ipiv, split16, LU_NB, base, pivoted, l21, ldl);
    } else {
    bool side_copy = false;   // an L21 copy was enqueued on the side stream (ev_copied follows the latest)
    int k = 0, jb = jb0;
    while (k < N) {
        BH_HIP(hipStreamWaitEvent(main.s, ev_panel, 0));   // panel k factored
        // the row swaps below reach into the previous panel's L21
        if (pivoted && side_copy) BH_HIP(hipStreamWaitEvent(main.s, ev_copied, 0));
        if (pivoted) laswp(main, k, A + k, lda, ipiv + k, 0, jb, k);
        const int rest = N - k - jb;
        if (rest <= 0) break;
        T* A12 = A + (size_t)(k + jb) * lda + k;
        if (pivoted) laswp(main, rest, A12, lda, ipiv + k, 0, jb, k);
        const int jb2 = width(rest);
        const T* W = ctx->lu.winv_l[p].as<T>();
        T* U = ctx->lu.u12.as<T>();   // U12 = L11^-1 A12, jb x rest, ld jb
        if (split16) {
            k_fill_u32<<<1, 64, 0, main.s>>>(ctx->split_mx.as<unsigned int>(), 2, 0u);
            BH_HIP(hipGetLastError());
        }
        // the next panel's columns first, then hand them to the side stream
        bool f16 = false;
        if constexpr (std::is_same<T, float>::value) {
            if (ctx->lu16) {
                lu16_cols(ctx, main, jb, rest, 0, jb2, A12, lda, p);
                f16 = true;
            }
        }
        if (!f16) {
            gemm_nn(main.h, jb, jb2, jb, W, jb, A12, lda, U, jb);
            if constexpr (std::is_same<T, float>::value)
                if (split16) split_b(ctx, main.s, jb, jb2, U, jb, 0, 0);
            trailing_update<T>(ctx, main.s, main.h, rest, jb2, jb, l21, ldl, U, jb, A12 + jb, lda, split16, 0, p, 0);
        }
        // the panel's small kernels only hide well behind a long update; a
        // short one is cheaper to run first and then factor the panel alone
        const bool overlap = ctx->lookahead && rest - jb2 >= ctx->la_min_rest;
        const Lane pl = overlap ? side : main;
        if (overlap) BH_HIP(hipEventRecord(ev_ready, main.s));
        // the rest of the step is enqueued before the next panel's ~70 small
        // launches: enqueued after them, it would reach the device only once
        // the host had issued them all, and the main stream would idle meanwhile
        if (rest > jb2) {
            T* Uw = U + (size_t)jb2 * jb;
            if constexpr (std::is_same<T, float>::value)
                if (f16) lu16_cols(ctx, main, jb, rest, jb2, rest - jb2, A12, lda, p);
            if (!f16) {
                gemm_nn(main.h, jb, rest - jb2, jb, W, jb, A12 + (size_t)jb2 * lda, lda, Uw, jb);
                if constexpr (std::is_same<T, float>::value)
                    if (split16) split_b(ctx, main.s, jb, rest - jb2, Uw, jb, jb2, 1);
                trailing_update<T>(ctx, main.s, main.h, rest, rest - jb2, jb, l21, ldl, Uw, jb,
                                   A12 + (size_t)jb2 * lda + jb, lda, split16, jb2, p, 1);
            }
        }
        if (!f16) copy2d(main.s, jb, rest, U, jb, A12, lda);   // U12 into the factor (lu16_cols wrote it)
        if (overlap) BH_HIP(hipStreamWaitEvent(side.s, ev_ready, 0));
        panel_start(ctx, pl, rest, jb2, A12 + jb, lda, ipiv + k + jb, k + jb, base, p ^ 1, split16);
        const T* l21n = nullptr;
        int ldln = 1;
        pivoted = panel_finish(ctx, pl, rest, jb2, A12 + jb, lda, ipiv + k + jb, k + jb, base, p ^ 1, split16, &l21n,
                               &ldln);
        BH_HIP(hipEventRecord(ev_panel, pl.s));
        // an accepted panel's L21 into the factor, after the hand-over (the
        // main stream reads it from the step buffers); the next panel's
        // hand-over event on this lane comes after the copy
        if (!pivoted) {
            copy2d(pl.s, rest - jb2, jb2, l21n, ldln, A12 + jb + jb2, lda);
            if (pl.s != main.s) {
                BH_HIP(hipEventRecord(ev_copied, pl.s));
                side_copy = true;
            }
        }
        k += jb;
        jb = jb2;
        p ^= 1;
        l21 = l21n;
        ldl = ldln;
    }
    BH_HIP(hipStreamWaitEvent(main.s, ev_panel, 0));
    if (side_copy) BH_HIP(hipStreamWaitEvent(main.s, ev_copied, 0));
    }
    }
    unsigned int flags[3] = {0, 0, 0};
    BH_HIP(hipMemcpyAsync(flags, ws, 12, hipMemcpyDeviceToHost, main.s));
    std::vector<unsigned int> blocks(ctx->split_fixed_on ? ctx->split_calls : 0);
    if (!blocks.empty())
        BH_HIP(hipMemcpyAsync(blocks.data(), ctx->split_log.p, sizeof(unsigned int) * blocks.size(),
                              hipMemcpyDeviceToHost, main.s));
    BH_HIP(hipStreamSynchronize(main.s));
    collect_gemm_timing(ctx);
    for (unsigned int b : blocks) {
        float m;
        std::memcpy(&m, &b, sizeof(m));
        const float sc = std::ldexp(m, ctx->split_eb);
        if (!(m == 0.0f || (sc >= 0x1p-8f && sc <= 32768.0f))) ctx->split_redo = true;   // NaN fails too
    }
    ctx->split_fixed_on = false;
    if (f16) {
        unsigned int mb = 0;
        BH_HIP(hipMemcpyAsync(&mb, ctx->lu.f16log.p, sizeof(mb), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        float m;
        std::memcpy(&m, &mb, sizeof(m));
        ctx->lu16_overflow = !(m <= 32768.0f);
    }
    BH_CHECK(!(flags[1] & 1u), BEMHIP_ERR_HIP, "panel grid barrier timed out");
    if (flags[1] & 2u) return (int)flags[2];  // 1-based first zero pivot
    return 0;
}

template int lu_factor<double>(Solver*, int, double*, int*, bool);
template int lu_factor<float>(Solver*, int, float*, int*, bool);

// --------------------------------------------------------------- solves --
constexpr int TS = 128;  // triangular solve block

template <class T, bool LOWER>
__global__ __launch_bounds__(TS) void k_trsv_diag(int N, const T* __restrict__ A, double* __restrict__ x, int j0,
                                                  int bs) {
    __shared__ double xs[TS];
    const int t = threadIdx.x;
    if (t < bs) xs[t] = x[j0 + t];
    __syncthreads();
    if (LOWER) {
        for (int c = 0; c < bs; ++c) {
            const double xc = xs[c];
            if (t > c && t < bs) xs[t] -= (double)A[(size_t)(j0 + c) * N + j0 + t] * xc;
            __syncthreads();
        }
    } else {
        for (int c = bs - 1; c >= 0; --c) {
            if (t == c) xs[c] /= (double)A[(size_t)(j0 + c) * N + j0 + c];
            __syncthreads();
            const double xc = xs[c];
            if (t < c) xs[t] -= (double)A[(size_t)(j0 + c) * N + j0 + t] * xc;
            __syncthreads();
        }
    }
    if (t < bs) x[j0 + t] = xs[t];
}

// rows [r0, r1): x[r] -= sum_c A[r, j0 + c] x[j0 + c]
template <class T>
__global__ __launch_bounds__(256) void k_trsv_update(int N, const T* __restrict__ A, double* __restrict__ x, int j0,
                                                     int bs, int r0, int r1) {
    __shared__ double xb[TS];
    if (threadIdx.x < bs) xb[threadIdx.x] = x[j0 + threadIdx.x];
    __syncthreads();
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    double acc = 0.0;
    const T* a = A + (size_t)j0 * N + r;
    for (int c = 0; c < bs; ++c) acc += (double)a[(size_t)c * N] * xb[c];
    x[r] -= acc;
}

__global__ void k_gather(int n, const int* __restrict__ perm, const double* __restrict__ src, double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// pivots (absolute, sequential) -> gather permutation: (P b)[i] = b[perm[i]]
void pivots_to_perm(const std::vector<int>& ipiv, std::vector<int>& perm) {
    const int n = (int)ipiv.size();
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        // partial pivoting only ever swaps row k with a row at or below it
        BH_CHECK(ipiv[k] >= k && ipiv[k] < n, BEMHIP_ERR_HIP,
                 "LU pivot " + std::to_string(ipiv[k]) + " of column " + std::to_string(k) + " out of range");
        std::swap(perm[k], perm[ipiv[k]]);
    }
}

// ---- persistent triangular solves -------------------------------------
// FP64 inverses of the TS x TS diagonal blocks (column-major per block), one
// work-group per block (grid.y: problem of a batch, A and inv sA, sInv
// elements further).  The block is staged in LDS with coalesced loads; thread
// (i, h) holds half h of row i of the inverse in registers, and step k
// broadcasts the finished row k through LDS to the rows it updates
// (right-looking substitution, one barrier a step, the broadcast row
// double-buffered; a half that is still zero in row k is skipped):
//   lower, unit:  X[i,:] = e_i - sum_{k<i} L[i][k] X[k,:]
//   upper:        X[i,:] = (e_i - sum_{k>i} U[i][k] X[k,:]) / U[i][i]
// Rows and columns beyond a short last block are the identity's.
template <class T, bool LOWER>
__global__ __launch_bounds__(2 * TS) void k_diag_inv(int N, const T* __restrict__ A, double* __restrict__ inv,
                                                     long long sA, long long sInv) {
    constexpr int HW = TS / 2;
    __shared__ T s[TS][TS + 1];     // s[i][k] = A(j0 + i, j0 + k)
    __shared__ double rb[2][TS];    // the broadcast row, by step parity
    A += blockIdx.y * sA;
    inv += blockIdx.y * sInv;
    const int b = blockIdx.x, t = threadIdx.x;
    const int i = t & (TS - 1), h = t >> 7;   // this thread's row and half (a wave-uniform half)
    const int j0 = b * TS, bs = min(TS, N - j0);
    const T* Ab = A + (size_t)j0 * N + j0;
    for (int c = h; c < TS; c += 2) s[i][c] = (i < bs && c < bs) ? Ab[(size_t)c * N + i] : (i == c ? T(1) : T(0));
    double x[HW];
#pragma unroll
    for (int c = 0; c < HW; ++c) x[c] = (h * HW + c == i) ? 1.0 : 0.0;
    __syncthreads();
#pragma unroll 1
    for (int q = 0; q < TS; ++q) {
        const int k = LOWER ? q : TS - 1 - q;
        if (i == k) {
            if (!LOWER) {
                const double r = 1.0 / (double)s[k][k];
#pragma unroll
                for (int c = 0; c < HW; ++c) x[c] *= r;
            }
#pragma unroll
            for (int c = 0; c < HW; ++c) rb[q & 1][h * HW + c] = x[c];
        }
        __syncthreads();
        const bool below = LOWER ? i > k : i < k;
        // row k is zero left of column k (upper) / right of it (lower)
        const bool nz = LOWER ? h * HW <= k : h * HW + HW - 1 >= k;
        if (below && nz) {
            const double f = (double)s[i][k];
#pragma unroll
            for (int c = 0; c < HW; ++c) x[c] -= f * rb[q & 1][h * HW + c];
        }
    }
    double* o = inv + (size_t)b * TS * TS;
#pragma unroll
    for (int c = 0; c < HW; ++c) o[(size_t)(h * HW + c) * TS + i] = x[c];
}

// One work-group per TS-row block, taken in dependency order through a ticket
// (so every block a work-group waits on belongs to an already running one).
// b holds the right-hand side, x receives the solution and holds the marker
// TRSV_PENDING in every entry on entry.  Block i accumulates A_ij x_j over the
// blocks j it depends on as they are published: each round the four waves
// read the next TRSV_POLL blocks' entries of x (readiness and values in one
// round trip: a block is published once none of its entries is the marker),
// then the work-group consumes the ready prefix.  Finally
// x_i = inv_i (b_i - sum) is published with agent-scope stores.
constexpr int TRSV_POLL = 8;
constexpr unsigned long long TRSV_PENDING = 0x7FF4DEADBEEF0000ull;   // a NaN payload no arithmetic produces

// x[q n + i] = the pending marker for every problem q with active[q] != 0
// (active null: all) of a batch of problems of n entries
__global__ void k_fill_pending(int n, int cnt, double* __restrict__ x, const int* __restrict__ active) {
    const long long tot = (long long)n * cnt;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x)
        if (!active || active[i / n]) x[i] = __longlong_as_double((long long)TRSV_PENDING);
}
static void fill_pending(hipStream_t st, int n, int cnt, double* x, const int* active) {
    const long long tot = (long long)n * cnt;
    k_fill_pending<<<(int)std::min<long long>(4096, (tot + 255) / 256), 256, 0, st>>>(n, cnt, x, active);
    BH_HIP(hipGetLastError());
}

// Batched (tb.gticket set): the solves of a batch of problems in one launch;
// the ticket names the problem (ticket / blocks) and its block in solve order,
// problem q's factor, inverses, b and x (ld N) and sync words start q sA,
// q sInv, q N, q N, q sstride elements further, and the problems with
// active[q] == 0 are skipped (all their blocks return at once, so nothing
// waits for them).
struct TrsvBat {
    unsigned int* gticket;
    long long sA, sInv;
    int sstride;
    const int* active;
};
// TRANS: the system is the transpose of the triangle named by LOWER's
// counterpart: LOWER solves U^T x = b (U^T is lower), !LOWER solves L^T x = b,
// both on the factor stored in A, with inv holding the diagonal-block inverses
// of U (resp. L) themselves (the inverse of a transposed block is the
// transposed inverse).  Row i of U^T is column i of U, contiguous in A: a
// half-wave reads 16 consecutive entries of each of 4 such rows per load and
// keeps per-row partial sums, reduced over the 16 lanes once at the end.
template <class T, bool LOWER, bool TRANS = false>
__global__ __launch_bounds__(256) void k_trsv_persist(int N, const T* __restrict__ A, const double* __restrict__ inv,
                                                      const double* __restrict__ b, double* __restrict__ x,
                                                      unsigned int* __restrict__ sync, TrsvBat tb) {
    __shared__ double xs[TRSV_POLL][TS];
    __shared__ double part[2][TS];
    __shared__ int s_t, s_ready, s_ok[TRSV_POLL];
    const int nb = (N + TS - 1) / TS;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = (int)atomicAdd(tb.gticket ? tb.gticket : sync, 1u);
    __syncthreads();
    int t = s_t;
    if (tb.gticket) {
        const int q = t / nb;
        t -= q * nb;
        if (tb.active && !tb.active[q]) return;   // the same for the whole work-group
        A += q * tb.sA;
        inv += q * tb.sInv;
        b += (long long)q * N;
        x += (long long)q * N;
        sync += (long long)q * tb.sstride;
    }
    unsigned int* err = sync + 1;      // this launch: a work-group gave up waiting
    unsigned int* sticky = sync + 2;   // any launch since the host cleared it
    const int ib = LOWER ? t : nb - 1 - t;
    const int i0 = ib * TS, bs = min(TS, N - i0);
    const int r = tid & (TS - 1), h = tid >> 7;
    const int wv = tid >> 6, ln = tid & 63;
    // the diagonal-block inverse does not depend on x: load this thread's half
    // row of it now so the final product is off the dependency chain
    // (held in the factor's precision: an FP32 factor only preconditions the
    // FP64 refinement; with the prefetched tile below the FP32 kernel runs one
    // work-group per compute unit, which measured faster than two with spills)
    T ivr[TS / 2];
    if constexpr (TRANS) {
        const double* iv = inv + (size_t)ib * TS * TS + (size_t)r * TS + h * (TS / 2);
#pragma unroll
        for (int c = 0; c < TS / 2; ++c) ivr[c] = (T)iv[c];
    } else {
        const double* iv = inv + (size_t)ib * TS * TS + (size_t)(h * (TS / 2)) * TS + r;
#pragma unroll
        for (int c = 0; c < TS / 2; ++c) ivr[c] = (T)iv[(size_t)c * TS];
    }
    // TRANS: lane group g4 of wave wv holds rows wv * 32 + 4 rg + g4 (rg < 8)
    // and columns u16 + 16 k (k < 8) of every block it consumes
    const int g4 = ln >> 4, u16 = ln & 15;
    // FP32: the tile of the last block this one depends on: once that block
    // is published its product is all that is left on the dependency chain,
    // so its half row is loaded now as well (FP64 lacks the registers)
    const int tl = t - 1;                      // that block's position in the solve order
    const int jl = LOWER ? tl : nb - 1 - tl;   // and its index
    const bool pre = sizeof(T) == 4 && t > 0 && (TRANS || r < bs) && N - jl * TS >= TS;
    T apre[TS / 2];
    if (pre) {
        if constexpr (TRANS) {
#pragma unroll
            for (int rg = 0; rg < 8; ++rg) {
                const int rr = wv * 32 + 4 * rg + g4;
                const T* row = A + (size_t)(i0 + rr) * N + jl * TS + u16;
#pragma unroll
                for (int k = 0; k < 8; ++k) apre[8 * rg + k] = rr < bs ? row[16 * k] : (T)0;
            }
        } else {
            const T* col = A + (size_t)(jl * TS + h * (TS / 2)) * N + i0 + r;
#pragma unroll
            for (int c = 0; c < TS / 2; ++c) apre[c] = col[(size_t)c * N];
        }
    }
    double acc = 0.0;
    double acct[8] = {};   // TRANS: the partial sums of rows wv * 32 + 4 rg + g4
    unsigned int spins = 0;
    for (int q = 0; q < t;) {
        // wave wv reads blocks q + wv and q + 4 + wv of x, two entries a lane
#pragma unroll
        for (int k = 0; k < TRSV_POLL / 4; ++k) {
            const int slot = wv + 4 * k, pb = q + slot;
            bool ok = pb < t;
            if (ok) {
                const int jb = LOWER ? pb : nb - 1 - pb;
                const int c = jb * TS + 2 * ln;
                double v0 = 0.0, v1 = 0.0;
                if (c < N) v0 = __hip_atomic_load(&x[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (c + 1 < N) v1 = __hip_atomic_load(&x[c + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = (unsigned long long)__double_as_longlong(v0) != TRSV_PENDING &&
                     (unsigned long long)__double_as_longlong(v1) != TRSV_PENDING;
                xs[slot][2 * ln] = v0;
                xs[slot][2 * ln + 1] = v1;
            }
            ok = __ballot(!ok) == 0ull;
            if (ln == 0) s_ok[slot] = ok;
        }
        // the error word, every round: once one work-group has given up, the
        // others stop waiting as well
        const unsigned int e = tid == 0 ? __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        __syncthreads();
        if (tid == 0) {
            int nready = 0;
            while (nready < TRSV_POLL && q + nready < t && s_ok[nready]) ++nready;
            if (nready == 0 && !e) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 24))
                    __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    nready = -1;   // poll again
            }
            if (nready == 0 || e) {
                // give up waiting: consume the polled blocks as they stand and
                // finish (results invalid; the host reads the sticky word)
                __hip_atomic_fetch_or(sticky, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                nready = min(t - q, TRSV_POLL);
            }
            s_ready = max(nready, 0);
        }
        __syncthreads();
        const int nr = s_ready;
        if constexpr (TRANS) {
            for (int bb = 0; bb < nr; ++bb) {
                double xv[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) xv[k] = xs[bb][u16 + 16 * k];
                if (pre && q + bb == tl) {
#pragma unroll
                    for (int rg = 0; rg < 8; ++rg)
#pragma unroll
                        for (int k = 0; k < 8; ++k) acct[rg] += (double)apre[8 * rg + k] * xv[k];
                    continue;
                }
                const int jb = LOWER ? q + bb : nb - 1 - (q + bb);
                const int j0 = jb * TS, js = min(TS, N - j0);
#pragma unroll
                for (int rg = 0; rg < 8; ++rg) {
                    const int rr = wv * 32 + 4 * rg + g4;
                    if (rr >= bs) continue;
                    const T* row = A + (size_t)(i0 + rr) * N + j0 + u16;
                    if (js == TS) {
                        T v[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) v[k] = row[16 * k];
#pragma unroll
                        for (int k = 0; k < 8; ++k) acct[rg] += (double)v[k] * xv[k];
                    } else {
                        for (int k = 0; k < 8; ++k)
                            if (u16 + 16 * k < js) acct[rg] += (double)row[16 * k] * xv[k];
                    }
                }
            }
        } else if (r < bs) {
            for (int bb = 0; bb < nr; ++bb) {
                if (pre && q + bb == tl) {
#pragma unroll
                    for (int c = 0; c < TS / 2; ++c) acc += (double)apre[c] * xs[bb][h * (TS / 2) + c];
                    continue;
                }
                const int jb = LOWER ? q + bb : nb - 1 - (q + bb);
                const int j0 = jb * TS, js = min(TS, N - j0);
                const int c0 = h * (TS / 2), c1 = min(js, c0 + TS / 2);
                const T* col = A + (size_t)(j0 + c0) * N + i0 + r;
                if (c1 - c0 == TS / 2) {
#pragma unroll
                    for (int cc = 0; cc < TS / 2; cc += 16) {
                        T v[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u) v[u] = col[(size_t)(cc + u) * N];
#pragma unroll
                        for (int u = 0; u < 16; ++u) acc += (double)v[u] * xs[bb][c0 + cc + u];
                    }
                } else {
                    for (int c = c0; c < c1; ++c, col += N) acc += (double)col[0] * xs[bb][c];
                }
            }
        }
        q += nr;
        __syncthreads();   // xs and s_ok are refilled by the next round
    }
    if constexpr (TRANS) {
        // sum each row's partials over its 16 lanes (xor 1..8 stays in the group)
#pragma unroll
        for (int rg = 0; rg < 8; ++rg) {
            double v = acct[rg];
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (u16 == 0) {
                part[0][wv * 32 + 4 * rg + g4] = v;
                part[1][wv * 32 + 4 * rg + g4] = 0.0;
            }
        }
    } else {
        part[h][r] = acc;
    }
    __syncthreads();
    if (tid < TS) xs[0][tid] = tid < bs ? b[i0 + tid] - (part[0][tid] + part[1][tid]) : 0.0;
    __syncthreads();
    {
        // x_i = inv_i s: thread (r, h) sums its half of row r (entries beyond bs are
        // zero in s, and the inverse is the identity there)
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < TS / 2; ++c) v += ivr[c] * xs[0][h * (TS / 2) + c];
        part[h][r] = v;
    }
    __syncthreads();
    // publishing a block is storing its entries (each replaces the marker)
    if (tid < bs)
        __hip_atomic_store(&x[i0 + tid], part[0][tid] + part[1][tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T> void lu_prepare_solve(Solver* ctx, int N, const T* LU) {
    const int nb = (N + TS - 1) / TS;
    double* il = static_cast<double*>(ctx->lu.inv_lo.ensure(sizeof(double) * (size_t)nb * TS * TS));
    double* iu = static_cast<double*>(ctx->lu.inv_up.ensure(sizeof(double) * (size_t)nb * TS * TS));
    k_diag_inv<T, true><<<nb, 2 * TS, 0, ctx->stream>>>(N, LU, il, 0, 0);
    k_diag_inv<T, false><<<nb, 2 * TS, 0, ctx->stream>>>(N, LU, iu, 0, 0);
    BH_HIP(hipGetLastError());
    const size_t fb = sizeof(unsigned int) * (16 + (size_t)nb);
    if (ctx->lu.flag.bytes < fb) {
        ctx->lu.flag.ensure(fb);
        BH_HIP(hipMemsetAsync(ctx->lu.flag.p, 0, fb, ctx->stream));
        ctx->trsv_epoch = 0;
    }
    ctx->lu.inv_n = N;
    ctx->lu.inv_src = LU;
}
template void lu_prepare_solve<double>(Solver*, int, const double*);
template void lu_prepare_solve<float>(Solver*, int, const float*);

// x = L^-1 b (lower) or U^-1 b; trans: x = U^-T b (lower) or L^-T b; x is
// overwritten (b and x distinct)
template <class T>
static void trsv_persist(Solver* ctx, int N, const T* LU, const double* b, double* x, bool lower, bool trans = false) {
    const int nb = (N + TS - 1) / TS;
    unsigned int* sync = ctx->lu.flag.as<unsigned int>();
    fill_pending(ctx->stream, N, 1, x, nullptr);
    k_fill_u32<<<1, 64, 0, ctx->stream>>>(sync, 2, 0u);   // ticket, error word
    if (ctx->trsv_force_giveup) k_fill_u32<<<1, 64, 0, ctx->stream>>>(sync + 1, 1, 1u);
    const TrsvBat one{nullptr, 0, 0, 0, nullptr};
    const double* il = ctx->lu.inv_lo.as<double>();
    const double* iu = ctx->lu.inv_up.as<double>();
    if (trans) {
        if (lower)
            k_trsv_persist<T, true, true><<<nb, 256, 0, ctx->stream>>>(N, LU, iu, b, x, sync, one);
        else
            k_trsv_persist<T, false, true><<<nb, 256, 0, ctx->stream>>>(N, LU, il, b, x, sync, one);
    } else if (lower) {
        k_trsv_persist<T, true><<<nb, 256, 0, ctx->stream>>>(N, LU, il, b, x, sync, one);
    } else {
        k_trsv_persist<T, false><<<nb, 256, 0, ctx->stream>>>(N, LU, iu, b, x, sync, one);
    }
    BH_HIP(hipGetLastError());
}

// the sticky give-up word of the persistent triangular solves: cleared before
// a solve's first launch, read back after its last
static void trsv_sticky_clear(Solver* ctx) {
    if (ctx->trsv_persistent && ctx->lu.flag.p)
        k_fill_u32<<<1, 64, 0, ctx->stream>>>(ctx->lu.flag.as<unsigned int>() + 2, 1, 0u);
}
static void trsv_sticky_check(Solver* ctx) {
    if (!ctx->trsv_persistent || !ctx->lu.flag.p) return;
    unsigned int* h = reinterpret_cast<unsigned int*>(ctx->host_scal + 40);
    BH_HIP(hipMemcpyAsync(h, ctx->lu.flag.as<unsigned int>() + 2, sizeof(unsigned int), hipMemcpyDeviceToHost,
                          ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    BH_CHECK(*h == 0u, BEMHIP_ERR_HIP, "persistent triangular solve gave up waiting for a block (result invalid)");
}

template <class T> void lu_solve(Solver* ctx, int N, const T* LU, const int* perm, double* x) {
    hipStream_t st = ctx->stream;
    double* tmp = static_cast<double*>(ctx->vec[7].ensure(sizeof(double) * N));
    BH_HIP(hipMemcpyAsync(tmp, x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    k_gather<<<(N + 255) / 256, 256, 0, st>>>(N, perm, tmp, x);
    if (ctx->trsv_persistent && ctx->lu.inv_n == N) {
        trsv_persist(ctx, N, LU, x, tmp, true);    // tmp = L^-1 P b
        trsv_persist(ctx, N, LU, tmp, x, false);   // x = U^-1 tmp
        return;
    }
    for (int j0 = 0; j0 < N; j0 += TS) {
        const int bs = std::min(TS, N - j0);
        k_trsv_diag<T, true><<<1, TS, 0, st>>>(N, LU, x, j0, bs);
        const int r0 = j0 + bs;
        if (r0 < N) k_trsv_update<T><<<(N - r0 + 255) / 256, 256, 0, st>>>(N, LU, x, j0, bs, r0, N);
    }
    const int nb = (N + TS - 1) / TS;
    for (int b = nb - 1; b >= 0; --b) {
        const int j0 = b * TS, bs = std::min(TS, N - j0);
        k_trsv_diag<T, false><<<1, TS, 0, st>>>(N, LU, x, j0, bs);
        if (j0 > 0) k_trsv_update<T><<<(j0 + 255) / 256, 256, 0, st>>>(N, LU, x, j0, bs, 0, j0);
    }
    BH_HIP(hipGetLastError());
}
template void lu_solve<double>(Solver*, int, const double*, const int*, double*);
template void lu_solve<float>(Solver*, int, const float*, const int*, double*);

// ------------------------------------------------------ gemv, reductions --
constexpr int GV_CHUNKS = 32;

// grid.z: problem q of a batch (A sA elements further, x and the partials
// q N and q gridDim.y N further); problems with active[q] == 0 are skipped
__global__ __launch_bounds__(256) void k_gemv_part(int N, const double* __restrict__ A, const double* __restrict__ x,
                                                   double* __restrict__ part, int chunk, long long sA,
                                                   const int* __restrict__ active) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = blockIdx.y * chunk, c1 = min(N, c0 + chunk);
    const int q = blockIdx.z;
    if (r >= N || (active && !active[q])) return;
    A += q * sA;
    x += (long long)q * N;
    part += (long long)q * gridDim.y * N;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = c0;
    const double* col = A + (size_t)c0 * N + r;
    for (; c + 8 <= c1; c += 8, col += 8 * (size_t)N) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)u * N];
        a0 += v[0] * x[c];     a1 += v[1] * x[c + 1];
        a2 += v[2] * x[c + 2]; a3 += v[3] * x[c + 3];
        a0 += v[4] * x[c + 4]; a1 += v[5] * x[c + 5];
        a2 += v[6] * x[c + 6]; a3 += v[7] * x[c + 7];
    }
    for (; c < c1; ++c, col += N) a0 += col[0] * x[c];
    part[(size_t)blockIdx.y * N + r] = (a0 + a1) + (a2 + a3);
}

// grid.y: problem q of a batch (vectors q N further)
__global__ __launch_bounds__(256) void k_gemv_finish(int N, int chunks, const double* __restrict__ part,
                                                     const double* __restrict__ b, double* __restrict__ r,
                                                     const int* __restrict__ active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= N || (active && !active[q])) return;
    part += (long long)q * chunks * N;
    if (b) b += (long long)q * N;
    r += (long long)q * N;
    double s = 0.0;
    for (int k = 0; k < chunks; ++k) s += part[(size_t)k * N + i];
    r[i] = b ? b[i] - s : s;
}

void gemv64(Solver* ctx, int N, const double* A, const double* x, const double* b, double* r) {
    const int chunks = std::min(GV_CHUNKS, N);
    const int chunk = (N + chunks - 1) / chunks;
    const int nch = (N + chunk - 1) / chunk;
    double* part = static_cast<double*>(ctx->gemv_part.ensure(sizeof(double) * (size_t)nch * N));
    k_gemv_part<<<dim3((N + 255) / 256, nch), 256, 0, ctx->stream>>>(N, A, x, part, chunk, 0, nullptr);
    k_gemv_finish<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, nch, part, b, r, nullptr);
    BH_HIP(hipGetLastError());
}

// deterministic two-pass reductions: op 0 = max |x|, 1 = dot(x, y), 2 = sum(x)
template <int OP>
__global__ __launch_bounds__(256) void k_reduce_part(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                     double* __restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (OP == 0) acc = fmax(acc, fabs(x[i]));
        else if (OP == 1) acc += x[i] * y[i];
        else acc += x[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = OP == 0 ? fmax(s[threadIdx.x], s[threadIdx.x + w]) : s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

template <int OP>
static double reduce(Solver* ctx, int n, const double* x, const double* y) {
    const int nb = 256;
    double* part = static_cast<double*>(ctx->scal.ensure(sizeof(double) * 1024));
    k_reduce_part<OP><<<nb, 256, 0, ctx->stream>>>(n, x, y, part);
    k_reduce_part<OP == 0 ? 0 : 2><<<1, 256, 0, ctx->stream>>>(nb, part, nullptr, part + 512);
    BH_HIP(hipMemcpyAsync(ctx->host_scal, part + 512, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->host_scal[0];
}
double dev_absmax(Solver* ctx, int n, const double* x) { return reduce<0>(ctx, n, x, nullptr); }
double dev_dot(Solver* ctx, int n, const double* x, const double* y) { return reduce<1>(ctx, n, x, y); }

// y = a*x + b*y (+ c*z when z)
__global__ __launch_bounds__(256) void k_axpby(int n, double a, const double* __restrict__ x, double b,
                                               double* __restrict__ y, double c, const double* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = a * x[i] + b * y[i];
    if (z) v += c * z[i];
    y[i] = v;
}
static void axpby(Solver* ctx, int n, double a, const double* x, double b, double* y, double c = 0.0,
                  const double* z = nullptr) {
    k_axpby<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, a, x, b, y, c, z);
}

__global__ __launch_bounds__(256) void k_bicg_setup(int N, const double* __restrict__ A, double* __restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double d = A[(size_t)i * N + i];
    inv[i] = d != 0.0 ? 1.0 / d : 1.0;
}
__global__ __launch_bounds__(256) void k_mul(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * b[i];
}

// Default refinement tolerance of the mixed solvers: stop once the correction
// is below IR_TOL |x|.  The error left is about that times the contraction of
// a sweep, measured at 3e-7 .. 3e-6 on sphere systems of N = 10242 and 40962,
// so two sweeps leave ~1e-16.  The FP64 residual itself is only accurate to
// ~1e-11 .. 1e-10 of |x| at N = 40962, so a tolerance at that floor (1e-10)
// costs a third sweep at random.
constexpr double IR_TOL = 1e-9;
// The loop also stops, before a sweep, once the FP64 residual meets
// |b - G x|_inf <= IR_RTOL |b|_inf.  After the first sweep the error is the
// FP32-class error times the contraction (~1e-10 of |x| at N = 40962, where
// the second sweep's correction measured <= 1.2e-10), so the residual test
// usually ends the loop one triangular-solve pair earlier than the
// correction test; the error left is at most cond(G) IR_RTOL relative.
constexpr double IR_RTOL = 1e-10;

// Eigen 3.4 bicgstab() with DiagonalPreconditioner, zero initial guess
static void bicgstab(Solver* ctx, int N, const double* A, const double* rhs, double* x, double tol, int maxit,
                     bemhip_solve_stats* st) {
    if (tol <= 0.0) tol = 2.220446049250313e-16;
    if (maxit <= 0) maxit = 2 * N;
    double* w[9];
    for (int k = 0; k < 7; ++k) w[k] = static_cast<double*>(ctx->vec[k].ensure(sizeof(double) * N));
    double* inv = w[0]; double* r = w[1]; double* r0 = w[2]; double* v = w[3]; double* p = w[4];
    double* y = w[5]; double* s = w[6];
    double* z = static_cast<double*>(ctx->vec[8].ensure(sizeof(double) * N));
    double* t = static_cast<double*>(ctx->vec[9].ensure(sizeof(double) * N));
    hipStream_t sm = ctx->stream;
    k_bicg_setup<<<(N + 255) / 256, 256, 0, sm>>>(N, A, inv);
    BH_HIP(hipMemsetAsync(x, 0, sizeof(double) * N, sm));
    BH_HIP(hipMemsetAsync(v, 0, sizeof(double) * N, sm));
    BH_HIP(hipMemsetAsync(p, 0, sizeof(double) * N, sm));
    BH_HIP(hipMemcpyAsync(r, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
    BH_HIP(hipMemcpyAsync(r0, rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
    double r0_sq = dev_dot(ctx, N, r0, r0);
    const double rhs_sq = dev_dot(ctx, N, rhs, rhs);
    int it = 0;
    if (rhs_sq == 0.0) {
        st->iterations = 0; st->converged = 1; st->residual = 0.0;
        return;
    }
    double rho = 1.0, alpha = 1.0, om = 1.0;
    const double tol2 = tol * tol * rhs_sq;
    const double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
    int restarts = 0;
    double rr = r0_sq;
    while (rr > tol2 && it < maxit) {
        const double rho_old = rho;
        rho = dev_dot(ctx, N, r0, r);
        if (std::fabs(rho) < eps2 * r0_sq) {
            gemv64(ctx, N, A, x, rhs, r);
            BH_HIP(hipMemcpyAsync(r0, r, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
            rho = r0_sq = dev_dot(ctx, N, r, r);
            if (restarts++ == 0) it = 0;
        }
        const double beta = (rho / rho_old) * (alpha / om);
        // p = r + beta (p - om v)
        axpby(ctx, N, -beta * om, v, beta, p, 1.0, r);
        k_mul<<<(N + 255) / 256, 256, 0, sm>>>(N, inv, p, y);
        gemv64(ctx, N, A, y, nullptr, v);
        alpha = rho / dev_dot(ctx, N, r0, v);
        // s = r - alpha v
        BH_HIP(hipMemcpyAsync(s, r, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        axpby(ctx, N, -alpha, v, 1.0, s);
        k_mul<<<(N + 255) / 256, 256, 0, sm>>>(N, inv, s, z);
        gemv64(ctx, N, A, z, nullptr, t);
        const double tt = dev_dot(ctx, N, t, t);
        om = tt > 0.0 ? dev_dot(ctx, N, t, s) / tt : 0.0;
        // x += alpha y + om z ; r = s - om t
        axpby(ctx, N, alpha, y, 1.0, x, om, z);
        BH_HIP(hipMemcpyAsync(r, s, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        axpby(ctx, N, -om, t, 1.0, r);
        ++it;
        rr = dev_dot(ctx, N, r, r);
    }
    BH_HIP(hipGetLastError());
    st->iterations = it;
    st->residual = std::sqrt(rr / rhs_sq);
    st->converged = rr <= tol2 ? 1 : 0;
}

// ------------------------------------------------------------ front end --
static int* make_perm(Solver* ctx, int N, const int* d_ipiv) {
    ctx->h_ipiv.resize(N);
    BH_HIP(hipMemcpyAsync(ctx->h_ipiv.data(), d_ipiv, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    pivots_to_perm(ctx->h_ipiv, ctx->h_perm);
    int* d = static_cast<int*>(ctx->perm.ensure(sizeof(int) * N));
    BH_HIP(hipMemcpyAsync(d, ctx->h_perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return d;
}

static float elapsed(Solver* ctx, hipEvent_t a, hipEvent_t b) {
    BH_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// FP32 copy of an FP64 matrix (n entries)
__global__ void k_d2f_mat(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

void solve_device(Solver* ctx, int N, double* G64, float* G32, const double* b, double* x, int solver,
                  double tol, int maxit, bemhip_solve_stats* st, bool report_residual) {
    hipStream_t sm = ctx->stream;
    int* ipiv = static_cast<int*>(ctx->lu.ipiv.ensure(sizeof(int) * N));
    ctx->gemm_ms = 0.0; ctx->gemm_flops = 0.0; ctx->gemm_launches = 0;
    ctx->panels = 0; ctx->panels_pivoted = 0;
    // the timing events, destroyed on every exit (the checks below throw)
    struct Events {
        hipEvent_t e[3] = {};
        ~Events() { for (hipEvent_t v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (hipEvent_t& v : ev.e) BH_HIP(hipEventCreate(&v));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    BH_HIP(hipEventRecord(e0, sm));
    st->iterations = 0; st->converged = 1; st->correction = 0.0; st->residual = 0.0;
    const bool small1 = ctx->small_single && solver != BEMHIP_SOLVE_BICGSTAB64 && N <= SMALL_MAX_N;
    if (small1) {
        // FP64 GEPP and both solves in one work-group, in place on G64
        char* w = static_cast<char*>(ctx->small1.ensure(256));
        const SmallJob job{0, N, 0};
        BH_HIP(hipMemcpyAsync(w, &job, sizeof(job), hipMemcpyHostToDevice, sm));
        small_solve_launch(sm, reinterpret_cast<const SmallJob*>(w), 1, G64, b, x, reinterpret_cast<int*>(w + 64));
        BH_HIP(hipEventRecord(e1, sm));
        int info = 0;
        BH_HIP(hipMemcpyAsync(&info, w + 64, sizeof(int), hipMemcpyDeviceToHost, sm));
        BH_HIP(hipStreamSynchronize(sm));
        BH_CHECK(info == 0, BEMHIP_ERR_SINGULAR, "zero pivot in column " + std::to_string(info));
    } else if (solver == BEMHIP_SOLVE_BICGSTAB64) {
        BH_HIP(hipEventRecord(e1, sm));
        bicgstab(ctx, N, G64, b, x, tol, maxit, st);
    } else if (solver == BEMHIP_SOLVE_LU64) {
        const int info = lu_factor<double>(ctx, N, G64, ipiv, false);
        BH_CHECK(info == 0, BEMHIP_ERR_SINGULAR, "zero pivot in column " + std::to_string(info));
        int* perm = make_perm(ctx, N, ipiv);
        if (ctx->trsv_persistent) lu_prepare_solve<double>(ctx, N, G64);
        trsv_sticky_clear(ctx);
        BH_HIP(hipEventRecord(e1, sm));
        BH_HIP(hipMemcpyAsync(x, b, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        lu_solve<double>(ctx, N, G64, perm, x);
        trsv_sticky_check(ctx);
    } else if (solver == BEMHIP_SOLVE_LU32_IR64 || solver == BEMHIP_SOLVE_LU16X3_IR64 ||
               solver == BEMHIP_SOLVE_LU16_IR64) {
        BH_CHECK(G32 != nullptr, BEMHIP_ERR_ARG, "mixed solve needs the FP32 copy of G");
        ctx->lu16 = solver == BEMHIP_SOLVE_LU16_IR64;
        const int panels0 = ctx->panels, pivoted0 = ctx->panels_pivoted;
        int info = lu_factor<float>(ctx, N, G32, ipiv, solver == BEMHIP_SOLVE_LU16X3_IR64);
        if (info == -2) {
            // merged fp16 steps met a panel that needs row interchanges: the
            // FP32 copy again from G, and the factorisation step by step (the
            // discarded pass's panels are not counted)
            ctx->panels = panels0;
            ctx->panels_pivoted = pivoted0;
            k_d2f_mat<<<4096, 256, 0, sm>>>((size_t)N * N, G64, G32);
            BH_HIP(hipGetLastError());
            const bool keep = ctx->lu16_merge;
            ctx->lu16_merge = false;
            info = lu_factor<float>(ctx, N, G32, ipiv, false);
            ctx->lu16_merge = keep;
            ctx->lu16_unmerged += 1;
        }
        ctx->lu16 = false;
        if (solver == BEMHIP_SOLVE_LU16_IR64 && ctx->lu16_overflow) {
            // an operand left the fp16 range: the FP32 copy again from G and
            // the factorisation with FP32 products
            k_d2f_mat<<<4096, 256, 0, sm>>>((size_t)N * N, G64, G32);
            BH_HIP(hipGetLastError());
            info = lu_factor<float>(ctx, N, G32, ipiv, false);
            ctx->lu16_redos += 1;
        }
        if (ctx->split_redo) {
            // a U12 block fell outside the fixed scale's window: the FP32 copy
            // again from G, and the factorisation with a scale per block
            k_d2f_mat<<<4096, 256, 0, sm>>>((size_t)N * N, G64, G32);
            BH_HIP(hipGetLastError());
            const bool keep = ctx->split_fixed;
            ctx->split_fixed = false;
            info = lu_factor<float>(ctx, N, G32, ipiv, true);
            ctx->split_fixed = keep;
            ctx->split_redos += 1;
        }
        BH_CHECK(info == 0, BEMHIP_ERR_SINGULAR, "zero pivot in column " + std::to_string(info));
        int* perm = make_perm(ctx, N, ipiv);
        if (ctx->trsv_persistent) lu_prepare_solve<float>(ctx, N, G32);
        trsv_sticky_clear(ctx);
        BH_HIP(hipEventRecord(e1, sm));
        if (tol <= 0.0) tol = IR_TOL;
        if (maxit <= 0) maxit = 30;
        BH_HIP(hipMemcpyAsync(x, b, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        lu_solve<float>(ctx, N, G32, perm, x);
        double* d = static_cast<double*>(ctx->vec[6].ensure(sizeof(double) * N));
        const double bn = dev_absmax(ctx, N, b);
        int it = 0;
        bool conv = false;
        double cn = 0.0;
        while (it < maxit) {
            gemv64(ctx, N, G64, x, b, d);   // d = b - G x  (FP64 residual)
            const double rn = dev_absmax(ctx, N, d);
            if (!std::isfinite(rn)) { cn = rn; break; }
            if (rn <= IR_RTOL * bn) { conv = true; break; }
            lu_solve<float>(ctx, N, G32, perm, d);
            axpby(ctx, N, 1.0, d, 1.0, x);
            ++it;
            const double dn = dev_absmax(ctx, N, d), xn = dev_absmax(ctx, N, x);
            cn = xn > 0.0 ? dn / xn : dn;
            if (!std::isfinite(cn)) break;
            if (cn <= tol) { conv = true; break; }
        }
        trsv_sticky_check(ctx);
        st->iterations = it;
        st->converged = conv ? 1 : 0;
        st->correction = cn;
    } else {
        throw Error(BEMHIP_ERR_ARG, "unknown solver " + std::to_string(solver));
    }
    BH_HIP(hipEventRecord(e2, sm));
    const float f_ms = elapsed(ctx, e0, e1), s_ms = elapsed(ctx, e1, e2);
    st->ms_factor = f_ms;
    st->ms_solve = s_ms;
    st->gemm_ms = ctx->gemm_ms;
    st->gemm_flops = ctx->gemm_flops;
    st->gemm_launches = ctx->gemm_launches;
    st->panels = ctx->panels;
    st->panels_pivoted = ctx->panels_pivoted;
    // final FP64 residual |b - G x| / |b| (not for LU64, whose G was overwritten)
    if (report_residual && solver != BEMHIP_SOLVE_LU64 && !small1) {
        double* r = static_cast<double*>(ctx->vec[5].ensure(sizeof(double) * N));
        gemv64(ctx, N, G64, x, b, r);
        const double rn = dev_dot(ctx, N, r, r), bn = dev_dot(ctx, N, b, b);
        st->residual = bn > 0.0 ? std::sqrt(rn / bn) : std::sqrt(rn);
    }
}

// ------------------------------------------------- transposed solves --
// G = P^T L U, so G^T x = b is U^T L^T (P x) = b: x = P^T (L^-T U^-T b).
__global__ void k_d2f_vec(int n, const double* __restrict__ a, float* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
template <class T>
__global__ void k_scatter_perm(int n, const int* __restrict__ perm, const T* __restrict__ src,
                               double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = (double)src[i];
}

// r = b - G^T x in FP64: entry i is column i of G (contiguous) against x, a
// wave per column; v2: 16-byte loads (N even, G and x 16-byte aligned)
__global__ __launch_bounds__(256) void k_gemv_t(int N, const double* __restrict__ G, const double* __restrict__ x,
                                               const double* __restrict__ b, double* __restrict__ r, int v2) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + w; i < N; i += gridDim.x * 4) {
        const double* col = G + (size_t)i * N;
        double s0 = 0.0, s1 = 0.0;
        if (v2) {
            const double2* c2 = reinterpret_cast<const double2*>(col);
            const double2* x2 = reinterpret_cast<const double2*>(x);
            const int n2 = N >> 1;
            int j = l;
            for (; j + 192 < n2; j += 256) {
                double2 a[4], y[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) a[u] = c2[j + 64 * u];
#pragma unroll
                for (int u = 0; u < 4; ++u) y[u] = x2[j + 64 * u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    s0 += a[u].x * y[u].x;
                    s1 += a[u].y * y[u].y;
                }
            }
            for (; j < n2; j += 64) {
                const double2 a = c2[j], y = x2[j];
                s0 += a.x * y.x;
                s1 += a.y * y.y;
            }
        } else {
            for (int j = l; j < N; j += 64) s0 += col[j] * x[j];
        }
        double s = s0 + s1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (l == 0) r[i] = b[i] - s;
    }
}
static void gemv_t_residual(Solver* ctx, int N, const double* G, const double* x, const double* b, double* r) {
    const int v2 = (N % 2 == 0) && ((reinterpret_cast<uintptr_t>(G) | reinterpret_cast<uintptr_t>(x)) % 16 == 0);
    k_gemv_t<<<std::min(4096, (N + 3) / 4), 256, 0, ctx->stream>>>(N, G, x, b, r, v2);
    BH_HIP(hipGetLastError());
}

// x = G^-T b with the persistent triangular solves on the factor whose
// diagonal-block inverses are current (FP64 vectors); false: not prepared
template <class T>
static bool lu_solve_t_persist(Solver* ctx, int N, const T* LU, const int* perm, const double* b, double* x) {
    if (!ctx->trsv_persistent || ctx->lu.inv_n != N || ctx->lu.inv_src != (const void*)LU) return false;
    double* y = static_cast<double*>(ctx->vec[8].ensure(sizeof(double) * N));
    double* z = static_cast<double*>(ctx->vec[9].ensure(sizeof(double) * N));
    trsv_persist(ctx, N, LU, b, y, true, true);    // y = U^-T b
    trsv_persist(ctx, N, LU, y, z, false, true);   // z = L^-T y
    k_scatter_perm<double><<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, perm, z, x);
    BH_HIP(hipGetLastError());
    return true;
}

// x = G^-T b from FP64 factors
static void lu_solve_t64(Solver* ctx, int N, const double* LU, const int* perm, const double* b, double* x) {
    if (lu_solve_t_persist(ctx, N, LU, perm, b, x)) return;
    hipStream_t st = ctx->stream;
    double* t = static_cast<double*>(ctx->vec[8].ensure(sizeof(double) * N));
    BH_HIP(hipMemcpyAsync(t, b, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    BH_BLAS(rocblas_dtrsv(ctx->blas, rocblas_fill_upper, rocblas_operation_transpose, rocblas_diagonal_non_unit, N,
                          LU, N, t, 1));
    BH_BLAS(rocblas_dtrsv(ctx->blas, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_unit, N, LU,
                          N, t, 1));
    k_scatter_perm<double><<<(N + 255) / 256, 256, 0, st>>>(N, perm, t, x);
    BH_HIP(hipGetLastError());
}

// x = G^-T b from FP32 factors (b rounded to FP32; refinement recovers it)
static void lu_solve_t32(Solver* ctx, int N, const float* LU, const int* perm, const double* b, double* x) {
    if (lu_solve_t_persist(ctx, N, LU, perm, b, x)) return;
    hipStream_t st = ctx->stream;
    float* t = static_cast<float*>(ctx->vec[9].ensure(sizeof(float) * N));
    k_d2f_vec<<<(N + 255) / 256, 256, 0, st>>>(N, b, t);
    BH_BLAS(rocblas_strsv(ctx->blas, rocblas_fill_upper, rocblas_operation_transpose, rocblas_diagonal_non_unit, N,
                          LU, N, t, 1));
    BH_BLAS(rocblas_strsv(ctx->blas, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_unit, N, LU,
                          N, t, 1));
    k_scatter_perm<float><<<(N + 255) / 256, 256, 0, st>>>(N, perm, t, x);
    BH_HIP(hipGetLastError());
}

void solve_transposed_device(Solver* ctx, int N, const double* G64, const float* G32, int solver,
                             const double* b, double* x, double tol, int maxit, bemhip_solve_stats* st) {
    const int* perm = ctx->perm.as<int>();
    BH_CHECK(perm != nullptr, BEMHIP_ERR_ARG, "transposed solve needs the factors of a previous LU solve");
    BH_CHECK(!(ctx->small_single && N <= SMALL_MAX_N), BEMHIP_ERR_ARG,
             "transposed solves need the blocked LU's factors (BEMHIP_SMALL_SINGLE routes this size elsewhere)");
    st->iterations = 0; st->converged = 1; st->correction = 0.0; st->residual = 0.0;
    trsv_sticky_clear(ctx);
    if (solver == BEMHIP_SOLVE_LU64) {
        lu_solve_t64(ctx, N, G64, perm, b, x);
        trsv_sticky_check(ctx);
        return;
    }
    BH_CHECK(solver == BEMHIP_SOLVE_LU32_IR64 || solver == BEMHIP_SOLVE_LU16X3_IR64 ||
                 solver == BEMHIP_SOLVE_LU16_IR64,
             BEMHIP_ERR_ARG, "transposed solve needs an LU-based solver");
    if (tol <= 0.0) tol = IR_TOL;
    if (maxit <= 0) maxit = 30;
    lu_solve_t32(ctx, N, G32, perm, b, x);
    double* r = static_cast<double*>(ctx->vec[6].ensure(sizeof(double) * N));
    double* d = static_cast<double*>(ctx->vec[5].ensure(sizeof(double) * N));
    int it = 0;
    bool conv = false;
    double cn = 0.0;
    while (it < maxit) {
        gemv_t_residual(ctx, N, G64, x, b, r);   // r = b - G^T x in FP64
        lu_solve_t32(ctx, N, G32, perm, r, d);
        axpby(ctx, N, 1.0, d, 1.0, x);
        ++it;
        const double dn = dev_absmax(ctx, N, d), xn = dev_absmax(ctx, N, x);
        cn = xn > 0.0 ? dn / xn : dn;
        if (!std::isfinite(cn)) break;
        if (cn <= tol) { conv = true; break; }
    }
    trsv_sticky_check(ctx);
    st->iterations = it;
    st->converged = conv ? 1 : 0;
    st->correction = cn;
}

// ------------------------------------------- batched fp16 LU (size classes) --
// The problems of one size class advance in lock step: every launch of a step
// (top blocks, triangular inverses, packs, the three block products) covers
// all of them, so a step's launches are as large as one big problem's, and no
// decision goes through the host.  Panels are factored without interchanges;
// max |l| over the whole factorisation is kept per problem and checked once at
// the end (<= 1: partial pivoting would have kept the diagonal at every column,
// so the factors are GEPP's); a problem that fails it is solved again on its
// own.  Scales are per problem: e[q] from max |diag A| as in the single fp16 LU.

ClassWork::~ClassWork() {
    for (hipEvent_t& v : ev)
        if (v) (void)hipEventDestroy(v);
}

// per problem: the scale exponent from max |diag A|, and the words the
// factorisation merges into cleared (top-block max |l| and error words, the L21
// max |l|, the fp16 log)
__global__ __launch_bounds__(256) void k_cls_prep(int n, const float* __restrict__ A, long long sA, int* __restrict__ e,
                                                  unsigned int* __restrict__ top_sync, int tss,
                                                  unsigned long long* __restrict__ maxl64,
                                                  unsigned int* __restrict__ f16log) {
    __shared__ float red[256];
    const int q = blockIdx.x, t = threadIdx.x;
    const float* Aq = A + q * sA;
    float m = 0.0f;
    for (int i = t; i < n; i += 256) {
        const float a = fabsf(Aq[(size_t)i * n + i]);
        m = (a <= 3.0e38f) ? fmaxf(m, a) : __builtin_inff();
    }
    red[t] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] = fmaxf(red[t], red[t + w]);
        __syncthreads();
    }
    if (t == 0) {
        const float dm = red[0];
        int ex = 0;
        if (dm > 0.0f && dm <= 3.0e38f) frexpf(dm, &ex);
        e[q] = max(-100, min(100, ex));
        unsigned int* s = top_sync + (long long)q * tss;
        s[0] = 0u;
        s[1] = 0u;
        s[3] = 0u;
        maxl64[q] = 0ull;
        f16log[q] = 0u;
    }
}

// word `word` of each of cnt regions (stride words apart) = v, and *ticket = 0
__global__ void k_cls_words(int cnt, unsigned int* __restrict__ base, int stride, int word, unsigned int v,
                            unsigned int* __restrict__ ticket) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < cnt) base[(long long)q * stride + word] = v;
    if (q == 0 && ticket) *ticket = 0u;
}

__global__ void k_copy_f64(size_t n, const double* __restrict__ a, double* __restrict__ b) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// dst[q n + i] = src[v0[q] + i]
__global__ void k_cls_gather(int n, const double* __restrict__ src, const int* __restrict__ v0,
                             double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i < n) dst[(long long)q * n + i] = src[v0[q] + i];
}
// x[v0[q] + i] = src[q n + i]
__global__ void k_cls_scatter(int n, const double* __restrict__ src, const int* __restrict__ v0,
                              double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i < n) x[v0[q] + i] = src[(long long)q * n + i];
}

// refinement state of a class (cnt problems): doubles bn, corr, then ints
// active, conv, iters
struct IrState {
    double* bn;
    double* corr;
    int* active;
    int* conv;
    int* iters;
};
__device__ inline double block_max256(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] = fmax(red[t], red[t + w]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}
// max |x| with NaN propagated (fmax would drop it)
__device__ inline double absmax_nan(double m, double x) {
    const double a = fabs(x);
    return (a != a || m != m) ? __longlong_as_double(0x7ff8000000000000ll) : fmax(m, a);
}
// one work-group per problem: |b|_inf, every problem refining
__global__ __launch_bounds__(256) void k_cls_ir_init(int n, const double* __restrict__ b, IrState s) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = absmax_nan(m, b[(long long)q * n + i]);
    m = block_max256(m, red);
    if (threadIdx.x == 0) {
        s.bn[q] = m;
        s.corr[q] = 0.0;
        s.active[q] = 1;
        s.conv[q] = 0;
        s.iters[q] = 0;
    }
}
// the residual test before a sweep: stop once |r|_inf <= rtol |b|_inf (or r
// is not finite: stopped, not converged)
__global__ __launch_bounds__(256) void k_cls_ir_res(int n, const double* __restrict__ r, double rtol, IrState s) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    if (!s.active[q]) return;   // the same for the whole work-group
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = absmax_nan(m, r[(long long)q * n + i]);
    m = block_max256(m, red);
    if (threadIdx.x == 0) {
        if (!(m <= 1.0e308)) s.active[q] = 0;
        else if (m <= rtol * s.bn[q]) { s.active[q] = 0; s.conv[q] = 1; }
    }
}
// after a sweep's solves: x += d, then stop once |d|_inf <= tol |x|_inf
__global__ __launch_bounds__(256) void k_cls_ir_corr(int n, double* __restrict__ x, const double* __restrict__ d,
                                                     double tol, IrState s) {
    __shared__ double red[256];
    const int q = blockIdx.x;
    if (!s.active[q]) return;
    double dm = 0.0, xm = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const long long k = (long long)q * n + i;
        const double dv = d[k], xv = x[k] + dv;
        x[k] = xv;
        dm = absmax_nan(dm, dv);
        xm = absmax_nan(xm, xv);
    }
    dm = block_max256(dm, red);
    xm = block_max256(xm, red);
    if (threadIdx.x == 0) {
        const double cn = xm > 0.0 ? dm / xm : dm;
        s.iters[q] += 1;
        s.corr[q] = cn;
        if (!(cn <= 1.0e308)) s.active[q] = 0;
        else if (cn <= tol) { s.active[q] = 0; s.conv[q] = 1; }
    }
}

// the device base of the class's publication stamps (tickets word CLS_EBASE)
// advances by CLS_EPOCH_STEP before every factorisation; a factorisation and
// its solves stamp fewer launches than that
constexpr int CLS_EBASE = 2;
constexpr unsigned int CLS_EPOCH_STEP = 1u << 16;
__global__ void k_epoch_advance(unsigned int* base) { *base += CLS_EPOCH_STEP; }

namespace {
constexpr int CLS_TRSV_WORDS_MIN = 16;
inline float p2(int k) { return std::ldexp(1.0f, k); }

struct ClassPtrs {
    ClassWork* w;
    float* A;          // problem 0's G32 block
    long long sA;      // elements between problems
    int n, cnt, NB;
    hipStream_t main, side;
    int max_wg;        // persistent update grid (0: one work-group per tile)
};

IrState ir_state(ClassWork* w) {
    double* d = w->scal.as<double>();
    int* i = reinterpret_cast<int*>(d + 2 * (size_t)w->cnt);
    return IrState{d, d + w->cnt, i, i + w->cnt, i + 2 * (size_t)w->cnt};
}

// the panel of the step at column k (m rows, jb columns) of every problem:
// top block, both triangular inverses, L21 = A21 U11^-1 in place with its
// tiles, L11^-1 as tiles (parity p)
void cls_panel(const ClassPtrs& c, hipStream_t s, int k, int m, int jb, int p, bool a21_ready) {
    ClassWork* w = c.w;
    const int n = c.n, cnt = c.cnt;
    float* Ap = c.A + (size_t)k * n + k;
    const int stripes = (jb + NPB - 1) / NPB;
    unsigned int* gt = w->tickets.as<unsigned int>();
    k_fill_u32<<<1, 64, 0, s>>>(gt, 1, 0u);
    k_np_top32<<<stripes * cnt, top_threads<float>(jb), 0, s>>>(
        jb, Ap, n, w->top_sync.as<unsigned int>(), ++w->top_epoch, w->linv.as<float>(), nullptr, gt, stripes, c.sA,
        TOP_SYNC_WORDS, (long long)TOP_STRIPES * NPB * NPB, w->tickets.as<unsigned int>() + CLS_EBASE);
    BH_CHECK(w->top_epoch < CLS_EPOCH_STEP, BEMHIP_ERR_ARG, "batched fp16 LU: too many top-block launches");
    BH_HIP(hipGetLastError());
    if (m <= jb) return;
    const long long sW = (long long)c.NB * c.NB;
    const size_t te = tri_inv_tmp_elems(jb);
    float* tinv = w->tinv.as<float>();
    const TriBat tb{cnt, c.sA, sW, (int64_t)(2 * te)};
    tri_inv2<float>(s, jb, Ap, n, w->winv_l[p].as<float>(), w->winv_u.as<float>(), tinv, tinv + te, &tb);
    unsigned int* log = w->f16log.as<unsigned int>();
    const int* e = w->e.as<int>();
    Mm16Bat bw{};
    bw.n = cnt; bw.sx = sW; bw.sa = (long long)w->tb_w; bw.e = e; bw.ek = 1; bw.ls = 1;
    mm16_pack(s, true, jb, jb, w->winv_u.as<float>(), jb, p2(-2), w->t_wu.p, log, nullptr, 0, nullptr, &bw);
    if (!a21_ready) {
        Mm16Bat ba{};
        ba.n = cnt; ba.sx = c.sA; ba.sa = (long long)w->tb_a21; ba.e = e; ba.ek = -1; ba.ls = 1;
        mm16_pack(s, false, m - jb, jb, Ap + jb, n, p2(8), w->t_a21.p, log, nullptr, 0, nullptr, &ba);
    }
    const Mm16Tiles tl{w->t_l21[p].p, false, jb, 0, 16384.0f, log, w->maxl64.as<unsigned long long>()};
    Mm16Bat bl{};
    bl.n = cnt; bl.sx = c.sA; bl.sa = (long long)w->tb_a21; bl.sb = (long long)w->tb_w; bl.st = (long long)w->tb_a21;
    bl.ls = 1;
    mm16_gemm(s, MM_STORE, 2, w->t_a21.p, jb, w->t_wu.p, jb, 0, m - jb, jb, jb, Ap + jb, n, p2(-6), 0, &tl, &bl);
    Mm16Bat bi{};
    bi.n = cnt; bi.sx = sW; bi.sa = (long long)w->tb_w; bi.ls = 1;
    mm16_pack(s, false, jb, jb, w->winv_l[p].as<float>(), jb, 1.0f, w->t_wl[p].p, log, nullptr, 0, nullptr, &bi);
}

// columns [c0, c0 + nc) of the block row of the step at column k (jb rows):
// U12 = L11^-1 A12 in place with its tiles, then the trailing update of those
// columns; upd: tiles the update writes of its result for the next step
void cls_cols(const ClassPtrs& c, int k, int jb, int rest, int c0, int nc, int p, bool a12_tiled,
              const Mm16Tiles* upd, size_t upd_stride) {
    if (nc <= 0) return;
    ClassWork* w = c.w;
    const int n = c.n, cnt = c.cnt;
    unsigned int* log = w->f16log.as<unsigned int>();
    const int* e = w->e.as<int>();
    float* a12 = c.A + (size_t)(k + jb + c0) * n + k;
    if (!a12_tiled) {
        const size_t off = (size_t)(c0 / 32) * (mm_pad(jb, MM_BK) / 16) * 1024;
        Mm16Bat bp{};
        bp.n = cnt; bp.sx = c.sA; bp.sa = (long long)w->tb_a12; bp.e = e; bp.ek = -1; bp.ls = 1;
        mm16_pack(c.main, true, nc, jb, a12, n, p2(8), w->t_a12.as<char>() + off, log, nullptr, 0, nullptr, &bp);
    }
    const Mm16Tiles tu{w->t_u12.p, true, jb, c0, p2(8), log, nullptr};
    Mm16Bat bu{};
    bu.n = cnt; bu.sx = c.sA; bu.sa = (long long)w->tb_w; bu.sb = (long long)w->tb_a12; bu.st = (long long)w->tb_a12;
    bu.e = e; bu.ek = 1; bu.ek2 = -1; bu.ls = 1;
    mm16_gemm(c.main, MM_STORE, 1, w->t_wl[p].p, jb, w->t_a12.p, jb, c0, jb, nc, jb, a12, n, p2(-8), 0, &tu, &bu);
    Mm16Bat bg{};
    bg.n = cnt; bg.sx = c.sA; bg.sa = (long long)w->tb_a21; bg.sb = (long long)w->tb_a12;
    bg.st = (long long)upd_stride; bg.e = e; bg.ek = 1; bg.ek2 = -1; bg.ls = 1;
    mm16_gemm(c.main, MM_SUB, 0, w->t_l21[p].p, jb, w->t_u12.p, jb, c0, rest, nc, jb, a12 + jb, n, p2(-22),
              c.max_wg, upd, &bg);
}

// the FP32 triangular solves of every problem (active: the refining ones) on x
// (ld n per problem), in place (through the class's fourth vector)
void cls_trsv(const ClassPtrs& c, double* x, const int* active) {
    ClassWork* w = c.w;
    const int n = c.n, cnt = c.cnt, nbk = w->nbk;
    const int tss = CLS_TRSV_WORDS_MIN + nbk;
    unsigned int* ts = w->tsync.as<unsigned int>();
    unsigned int* gt = w->tickets.as<unsigned int>() + 1;
    const long long sInv = (long long)nbk * TS * TS;
    double* tt = w->vecs.as<double>() + 3 * (size_t)cnt * n;
    for (int lower = 1; lower >= 0; --lower) {
        const double* in = lower ? x : tt;
        double* out = lower ? tt : x;
        fill_pending(c.main, n, cnt, out, active);
        k_cls_words<<<(cnt + 255) / 256, 256, 0, c.main>>>(cnt, ts, tss, 1, 0u, gt);
        const TrsvBat tb{gt, c.sA, sInv, tss, active};
        if (lower)
            k_trsv_persist<float, true><<<nbk * cnt, 256, 0, c.main>>>(n, c.A, w->inv_lo.as<double>(), in, out, ts,
                                                                      tb);
        else
            k_trsv_persist<float, false><<<nbk * cnt, 256, 0, c.main>>>(n, c.A, w->inv_up.as<double>(), in, out, ts,
                                                                       tb);
        BH_HIP(hipGetLastError());
    }
}

// one refinement sweep of the refining problems: r = b - G x (FP64), the
// residual test, d = (LU)^-1 r, x += d and the correction test
void cls_sweep(const ClassPtrs& c, const double* G64, const double* bb, double* xx, double* rr, double tol) {
    ClassWork* w = c.w;
    const int n = c.n, cnt = c.cnt, nch = w->nch;
    const int chunk = (n + nch - 1) / nch;
    const IrState s = ir_state(w);
    k_gemv_part<<<dim3((n + 255) / 256, nch, cnt), 256, 0, c.main>>>(n, G64, xx, w->part.as<double>(), chunk, c.sA,
                                                                     s.active);
    k_gemv_finish<<<dim3((n + 255) / 256, cnt), 256, 0, c.main>>>(n, nch, w->part.as<double>(), bb, rr, s.active);
    k_cls_ir_res<<<cnt, 256, 0, c.main>>>(n, rr, IR_RTOL, s);
    BH_HIP(hipGetLastError());
    cls_trsv(c, rr, s.active);
    k_cls_ir_corr<<<cnt, 256, 0, c.main>>>(n, xx, rr, tol, s);
    BH_HIP(hipGetLastError());
}

ClassPtrs class_ptrs(Solver* ctx, ClassWork* w, const ClassJob& cj, float* G32) {
    ClassPtrs c{};
    c.w = w;
    c.A = G32 + cj.gofs0;
    c.sA = gblock_stride(cj.n);
    c.n = cj.n;
    c.cnt = cj.cnt;
    c.NB = w->nb;
    c.main = ctx->stream;
    const bool la = ctx->lookahead && cj.lookahead;
    c.side = la ? ctx->stream2 : ctx->stream;
    c.max_wg = la && ctx->mm_reserve > 0 ? std::max(8, ctx->ncu - ctx->mm_reserve) : 0;
    return c;
}

// class vectors: b, x, r (cnt x n doubles each; a fourth is cls_trsv's)
double* cls_vec(ClassWork* w, int i) { return w->vecs.as<double>() + (size_t)i * w->cnt * w->n; }

void class_work_ensure(Solver* ctx, ClassWork* w, const ClassJob& cj) {
    const int n = cj.n, cnt = cj.cnt;
    const int NB = std::min(ctx->nb16, n);
    BH_CHECK(ctx->nb16 % MM_TILE == 0 && ctx->nb16 <= top_max_rows<float>(), BEMHIP_ERR_ARG,
             "batched fp16 LU: block must be a multiple of 256 and at most 1024");
    BH_CHECK(cnt >= 1 && (long long)cnt * TOP_STRIPES <= (1ll << 30), BEMHIP_ERR_ARG, "batched fp16 LU: bad class");
    const int nbk = (n + TS - 1) / TS;
    const bool fresh = w->n != n || w->cnt != cnt || w->nb != NB;
    w->n = n;
    w->cnt = cnt;
    w->nb = NB;
    w->nbk = nbk;
    {
        // column chunks of the residual GEMV, as gemv64 cuts them
        const int chunks = std::min(GV_CHUNKS, n), chunk = (n + chunks - 1) / chunks;
        w->nch = (n + chunk - 1) / chunk;
    }
    w->tb_a21 = mm_tile_bytes(n, NB);
    w->tb_w = mm_tile_bytes(NB, NB);
    w->tb_a12 = mm_tile_bytes(n + MM_TILE, NB);
    const size_t C = (size_t)cnt;
    w->e.ensure(sizeof(int) * C);
    w->tickets.ensure(256);
    w->maxl64.ensure(sizeof(unsigned long long) * C);
    w->f16log.ensure(sizeof(unsigned int) * C);
    w->linv.ensure(sizeof(float) * C * TOP_STRIPES * NPB * NPB);
    for (int p = 0; p < 2; ++p) {
        w->winv_l[p].ensure(sizeof(float) * C * NB * NB);
        w->t_l21[p].ensure(C * w->tb_a21);
        w->t_wl[p].ensure(C * w->tb_w);
    }
    w->winv_u.ensure(sizeof(float) * C * NB * NB);
    w->tinv.ensure(sizeof(float) * C * 2 * tri_inv_tmp_elems(NB));
    w->t_a21.ensure(C * w->tb_a21);
    w->t_wu.ensure(C * w->tb_w);
    w->t_a12.ensure(C * w->tb_a12);
    w->t_u12.ensure(C * w->tb_a12);
    w->inv_lo.ensure(sizeof(double) * C * nbk * TS * TS);
    w->inv_up.ensure(sizeof(double) * C * nbk * TS * TS);
    w->vecs.ensure(sizeof(double) * 4 * C * n);   // b, x, r and the solves' intermediate
    w->part.ensure(sizeof(double) * C * w->nch * n);
    w->scal.ensure((2 * sizeof(double) + 3 * sizeof(int)) * C);
    w->v0.ensure(sizeof(int) * C);
    const size_t tsw = C * TOP_SYNC_WORDS, trw = C * (CLS_TRSV_WORDS_MIN + nbk);
    if (fresh) w->v0_h.clear();
    if (fresh || w->top_sync.bytes < sizeof(unsigned int) * tsw || w->tsync.bytes < sizeof(unsigned int) * trw) {
        // flags start below every publication stamp
        w->top_sync.ensure(sizeof(unsigned int) * tsw);
        w->tsync.ensure(sizeof(unsigned int) * trw);
        k_fill_u32<<<(int)((tsw + 255) / 256), 256, 0, ctx->stream>>>(w->top_sync.as<unsigned int>(), (int)tsw, 0u);
        k_fill_u32<<<(int)((trw + 255) / 256), 256, 0, ctx->stream>>>(w->tsync.as<unsigned int>(), (int)trw, 0u);
        k_fill_u32<<<1, 64, 0, ctx->stream>>>(w->tickets.as<unsigned int>(), 64, 0u);
        BH_HIP(hipGetLastError());
        w->top_epoch = 0;
        w->trsv_epoch = 0;
    }
    if (!w->ev[0])
        for (hipEvent_t& v : w->ev) BH_HIP(hipEventCreate(&v));
}
}  // namespace

void class_lu16_prepare(Solver* ctx, ClassWork* w, const ClassJob& cj) {
    BH_CHECK(cj.n > 0 && cj.cnt > 0 && (int)cj.v0.size() == cj.cnt, BEMHIP_ERR_ARG, "batched fp16 LU: bad class");
    class_work_ensure(ctx, w, cj);
    if (w->v0_h != cj.v0) {
        // the problems' offsets into b and x, uploaded when they change
        BH_HIP(hipMemcpyAsync(w->v0.p, cj.v0.data(), sizeof(int) * cj.cnt, hipMemcpyHostToDevice, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        w->v0_h = cj.v0;
    }
}

void class_lu16_enqueue(Solver* ctx, ClassWork* w, const ClassJob& cj, double* G64, float* G32, const double* b,
                        double* x, double tol, int sweeps, bool events) {
    BH_CHECK(cj.n > 0 && cj.cnt > 0 && (int)cj.v0.size() == cj.cnt, BEMHIP_ERR_ARG, "batched fp16 LU: bad class");
    BH_CHECK(G32 != nullptr, BEMHIP_ERR_ARG, "batched fp16 LU needs the FP32 copy of G");
    if (tol <= 0.0) tol = IR_TOL;
    class_lu16_prepare(ctx, w, cj);
    const ClassPtrs c = class_ptrs(ctx, w, cj, G32);
    const int n = c.n, cnt = c.cnt, NB = c.NB;
    hipEvent_t ev_ready = w->ev[0], ev_panel = w->ev[1];
    if (events) BH_HIP(hipEventRecord(w->ev[2], c.main));
    w->timed = events;
    k_epoch_advance<<<1, 1, 0, c.main>>>(w->tickets.as<unsigned int>() + CLS_EBASE);
    w->top_epoch = 0;
    w->trsv_epoch = 0;
    k_cls_prep<<<cnt, 256, 0, c.main>>>(n, c.A, c.sA, w->e.as<int>(), w->top_sync.as<unsigned int>(), TOP_SYNC_WORDS,
                                        w->maxl64.as<unsigned long long>(), w->f16log.as<unsigned int>());
    BH_HIP(hipGetLastError());
    // factorisation: panel 0 on the main stream, then per step the next
    // panel's columns, the next panel on the side stream while the main stream
    // updates the rest (buffers of a panel's products alternate by parity)
    int k = 0, jb = std::min(NB, n), p = 0;
    cls_panel(c, c.main, 0, n, jb, 0, false);
    BH_HIP(hipEventRecord(ev_panel, c.main));
    bool a12_tiled = false;
    for (;;) {
        BH_HIP(hipStreamWaitEvent(c.main, ev_panel, 0));
        const int rest = n - k - jb;
        if (rest <= 0) break;
        const int jb2 = std::min(NB, rest);
        const bool a21 = rest > jb2;
        const Mm16Tiles ta{w->t_a21.p, false, jb2, 0, p2(8), w->f16log.as<unsigned int>(), nullptr, jb2};
        cls_cols(c, k, jb, rest, 0, jb2, p, a12_tiled, a21 ? &ta : nullptr, w->tb_a21);
        BH_HIP(hipEventRecord(ev_ready, c.main));
        if (c.side != c.main) BH_HIP(hipStreamWaitEvent(c.side, ev_ready, 0));
        cls_panel(c, c.side, k + jb, rest, jb2, p ^ 1, a21);
        BH_HIP(hipEventRecord(ev_panel, c.side));
        if (rest > jb2) {
            const Mm16Tiles tb{w->t_a12.p, true, jb2, 0, p2(8), w->f16log.as<unsigned int>(), nullptr};
            cls_cols(c, k, jb, rest, jb2, rest - jb2, p, a12_tiled, &tb, w->tb_a12);
            a12_tiled = true;
        } else {
            a12_tiled = false;
        }
        k += jb;
        jb = jb2;
        p ^= 1;
    }
    if (events) BH_HIP(hipEventRecord(w->ev[3], c.main));
    // solves: FP64 inverses of the 128 x 128 diagonal blocks, x0 = (LU)^-1 b,
    // then the refinement sweeps
    const long long sInv = (long long)w->nbk * TS * TS;
    k_diag_inv<float, true><<<dim3(w->nbk, cnt), 2 * TS, 0, c.main>>>(n, c.A, w->inv_lo.as<double>(), c.sA, sInv);
    k_diag_inv<float, false><<<dim3(w->nbk, cnt), 2 * TS, 0, c.main>>>(n, c.A, w->inv_up.as<double>(), c.sA, sInv);
    const int tss = CLS_TRSV_WORDS_MIN + w->nbk;
    k_cls_words<<<(cnt + 255) / 256, 256, 0, c.main>>>(cnt, w->tsync.as<unsigned int>(), tss, 2, 0u, nullptr);
    double* bb = cls_vec(w, 0);
    double* xx = cls_vec(w, 1);
    double* rr = cls_vec(w, 2);
    k_cls_gather<<<dim3((n + 255) / 256, cnt), 256, 0, c.main>>>(n, b, w->v0.as<int>(), bb);
    BH_HIP(hipGetLastError());
    k_copy_f64<<<std::min(4096, (int)(((size_t)cnt * n + 255) / 256)), 256, 0, c.main>>>((size_t)cnt * n, bb, xx);
    cls_trsv(c, xx, nullptr);
    k_cls_ir_init<<<cnt, 256, 0, c.main>>>(n, bb, ir_state(w));
    BH_HIP(hipGetLastError());
    for (int s = 0; s < sweeps; ++s) cls_sweep(c, G64 + cj.gofs0, bb, xx, rr, tol);
    k_cls_scatter<<<dim3((n + 255) / 256, cnt), 256, 0, c.main>>>(n, xx, w->v0.as<int>(), x);
    BH_HIP(hipGetLastError());
    if (events) BH_HIP(hipEventRecord(w->ev[4], c.main));
}

void class_lu16_finish(Solver* ctx, ClassWork* w, const ClassJob& cj, double* G64, float* G32, const double* b,
                       double* x, double tol, int maxit, bemhip_solve_stats* st) {
    if (tol <= 0.0) tol = IR_TOL;
    if (maxit <= 0) maxit = 30;
    const ClassPtrs c = class_ptrs(ctx, w, cj, G32);
    const int n = c.n, cnt = c.cnt;
    double* bb = cls_vec(w, 0);
    double* xx = cls_vec(w, 1);
    double* rr = cls_vec(w, 2);
    const int tss = CLS_TRSV_WORDS_MIN + w->nbk;
    std::vector<unsigned int> top((size_t)cnt * TOP_SYNC_WORDS), ts((size_t)cnt * tss), f16((size_t)cnt);
    std::vector<unsigned long long> ml((size_t)cnt);
    std::vector<char> sc(w->scal.bytes);
    auto read_status = [&] {
        BH_HIP(hipMemcpyAsync(sc.data(), w->scal.p, sc.size(), hipMemcpyDeviceToHost, c.main));
        BH_HIP(hipMemcpyAsync(ts.data(), w->tsync.p, sizeof(unsigned int) * ts.size(), hipMemcpyDeviceToHost, c.main));
        BH_HIP(hipStreamSynchronize(c.main));
    };
    BH_HIP(hipMemcpyAsync(top.data(), w->top_sync.p, sizeof(unsigned int) * top.size(), hipMemcpyDeviceToHost, c.main));
    BH_HIP(hipMemcpyAsync(ml.data(), w->maxl64.p, sizeof(unsigned long long) * cnt, hipMemcpyDeviceToHost, c.main));
    BH_HIP(hipMemcpyAsync(f16.data(), w->f16log.p, sizeof(unsigned int) * cnt, hipMemcpyDeviceToHost, c.main));
    read_status();
    float fms = 0.0f, sms = 0.0f;
    if (w->timed) {
        BH_HIP(hipEventElapsedTime(&fms, w->ev[2], w->ev[3]));
        BH_HIP(hipEventElapsedTime(&sms, w->ev[3], w->ev[4]));
    }
    w->ms_factor = fms;
    const double* sd = reinterpret_cast<const double*>(sc.data());
    auto sint = [&](int which, int q) { return reinterpret_cast<const int*>(sd + 2 * (size_t)cnt)[(size_t)which * cnt + q]; };
    // problems to solve again on their own: a multiplier above 1 (GEPP would
    // have interchanged rows), an fp16 overflow, a top block that gave up waiting
    std::vector<int> redo;
    int sweeps_done = 0;
    for (int q = 0; q < cnt; ++q) {
        const unsigned int* tq = top.data() + (size_t)q * TOP_SYNC_WORDS;
        unsigned long long tl = 0;
        std::memcpy(&tl, tq, sizeof(tl));
        double l1, l2;
        std::memcpy(&l1, &tl, sizeof(l1));
        std::memcpy(&l2, &ml[q], sizeof(l2));
        float fm;
        std::memcpy(&fm, &f16[q], sizeof(fm));
        BH_CHECK(tq[3] == 0u, BEMHIP_ERR_HIP, "batched panel top-block factorisation gave up waiting");
        if (!(l1 <= 1.0 && l2 <= 1.0 && fm <= 32768.0f) || q == ctx->batch_force_redo) redo.push_back(q);
        sweeps_done = std::max(sweeps_done, sint(2, q));
    }
    for (int q = 0; q < cnt; ++q)
        BH_CHECK(ts[(size_t)q * tss + 2] == 0u, BEMHIP_ERR_HIP,
                 "persistent triangular solve gave up waiting for a block (result invalid)");
    // more sweeps while a problem (not solved again) still refines
    int extra = 0;
    for (;;) {
        bool any = false;
        for (int q = 0; q < cnt; ++q)
            if (sint(0, q) && std::find(redo.begin(), redo.end(), q) == redo.end()) any = true;
        if (!any || sweeps_done + extra >= maxit) break;
        cls_sweep(c, G64 + cj.gofs0, bb, xx, rr, tol);
        ++extra;
        read_status();
        for (int q = 0; q < cnt; ++q)
            BH_CHECK(ts[(size_t)q * tss + 2] == 0u, BEMHIP_ERR_HIP,
                     "persistent triangular solve gave up waiting for a block (result invalid)");
    }
    if (extra > 0) {
        k_cls_scatter<<<dim3((n + 255) / 256, cnt), 256, 0, c.main>>>(n, xx, w->v0.as<int>(), x);
        BH_HIP(hipGetLastError());
    }
    bemhip_solve_stats sum{};
    sum.converged = 1;
    for (int q = 0; q < cnt; ++q) {
        if (std::find(redo.begin(), redo.end(), q) != redo.end()) continue;
        sum.iterations = std::max(sum.iterations, sint(2, q));
        sum.converged = sum.converged && sint(1, q);
        sum.correction = std::max(sum.correction, sd[cnt + q]);
    }
    sum.panels = cnt * ((n + c.NB - 1) / c.NB);
    for (int q : redo) {
        const size_t off = (size_t)cj.gofs0 + (size_t)q * c.sA;
        k_d2f_mat<<<4096, 256, 0, c.main>>>((size_t)n * n, G64 + off, G32 + off);
        BH_HIP(hipGetLastError());
        bemhip_solve_stats one{};
        solve_device(ctx, n, G64 + off, G32 + off, b + cj.v0[q], x + cj.v0[q], BEMHIP_SOLVE_LU16_IR64, tol, maxit,
                     &one, false);
        sum.iterations = std::max(sum.iterations, one.iterations);
        sum.converged = sum.converged && one.converged;
        sum.correction = std::max(sum.correction, one.correction);
        sum.panels += one.panels;
        sum.panels_pivoted += one.panels_pivoted;
        sum.ms_factor += one.ms_factor;
    }
    sum.ms_factor += fms;
    sum.ms_solve += sms;
    *st = sum;
}

}  // namespace bemhip

namespace bemhip {

void Solver::init(const Solver* cfg) {
    if (cfg) {
        la_min_rest = cfg->la_min_rest;
        lookahead = cfg->lookahead;
        speculate = cfg->speculate;
        np_top = cfg->np_top;
        np_top32 = cfg->np_top32;
        nb64 = cfg->nb64; nb32 = cfg->nb32; nb_split = cfg->nb_split; nb16 = cfg->nb16;
        trsv_persistent = cfg->trsv_persistent;
        trsv_force_giveup = cfg->trsv_force_giveup;
        lu16_merge = cfg->lu16_merge;
        lu16_bulk_first = cfg->lu16_bulk_first;
        batch_force_redo = cfg->batch_force_redo;
        mm_reserve = cfg->mm_reserve;
        timing = cfg->timing;
        split_fixed = cfg->split_fixed;
        small_single = cfg->small_single;
    }
    {
        int dev = 0;
        BH_HIP(hipGetDevice(&dev));
        BH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    BH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    BH_BLAS(rocblas_create_handle(&blas));
    BH_BLAS(rocblas_set_stream(blas, stream));
    BH_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    BH_HIP(hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking));
    for (auto& e : ev_m) BH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    BH_BLAS(rocblas_create_handle(&blas2));
    BH_BLAS(rocblas_set_stream(blas2, stream2));
    BH_HIP(hipHostMalloc((void**)&host_scal, 64 * sizeof(double), hipHostMallocDefault));
    for (auto& e : ev_la) BH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void Solver::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (stream3) (void)hipStreamSynchronize(stream3);
    for (auto& e : ev_m) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (auto& e : ev_la) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (auto& e : gemm_ev) if (e) (void)hipEventDestroy(e);
    gemm_ev.clear();
    if (blas2) { (void)rocblas_destroy_handle(blas2); blas2 = nullptr; }
    if (stream2) { (void)hipStreamDestroy(stream2); stream2 = nullptr; }
    if (stream3) { (void)hipStreamDestroy(stream3); stream3 = nullptr; }
    if (blas) { (void)rocblas_destroy_handle(blas); blas = nullptr; }
    if (host_scal) { (void)hipHostFree(host_scal); host_scal = nullptr; }
    if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
}

}  // namespace bemhip
