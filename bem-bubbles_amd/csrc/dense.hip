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
#include <vector>

#include "internal.hpp"

namespace bemhip {

constexpr int LU_NB = 256;   // outer block
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

// ---------------------------------------------------- split-fp16 update --
// max |x| over a column-major block (rows x cols, ld) -> bits of a float in *out
__global__ __launch_bounds__(256) void k_block_absmax(int rows, int cols, const float* __restrict__ X, int ld,
                                                      unsigned int* __restrict__ out) {
    float m = 0.0f;
    const size_t total = (size_t)rows * cols;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / rows, r = e % rows;
        m = fmaxf(m, fabsf(X[c * ld + r]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));  // non-negative floats order as uints
}

// scale exponents: A and B scaled so that max |.| maps to about 2^12
__device__ inline int scale_exp(unsigned int bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f)) return 0;
    int e;
    frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return 12 - e;
}

__global__ void k_split_alpha(const unsigned int* __restrict__ mx, float* __restrict__ alpha) {
    const int ea = scale_exp(mx[0]), eb = scale_exp(mx[1]);
    alpha[0] = -ldexpf(1.0f, -(ea + eb));
}

// Acat (rows x 3K, ld rows) = [hi | hi | lo] of 2^ea A
__global__ __launch_bounds__(256) void k_split_a(int rows, int K, const float* __restrict__ A, int lda,
                                                 const unsigned int* __restrict__ mx, __half* __restrict__ out) {
    const int ea = scale_exp(mx[0]);
    const size_t total = (size_t)rows * K;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / rows, r = e % rows;
        const float v = ldexpf(A[c * lda + r], ea);
        const __half h = __float2half_rn(v);
        const __half l = __float2half_rn(v - __half2float(h));
        out[c * rows + r] = h;
        out[(c + K) * rows + r] = h;
        out[(c + 2 * K) * rows + r] = l;
    }
}

// Bcat (3K x cols, ld 3K) = [hi; lo; hi] of 2^eb B
__global__ __launch_bounds__(256) void k_split_b(int K, int cols, const float* __restrict__ B, int ldb,
                                                 const unsigned int* __restrict__ mx, __half* __restrict__ out) {
    const int eb = scale_exp(mx[1]);
    const size_t total = (size_t)K * cols;
    const size_t K3 = 3 * (size_t)K;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / K, r = e % K;
        const float v = ldexpf(B[c * ldb + r], eb);
        const __half h = __float2half_rn(v);
        const __half l = __float2half_rn(v - __half2float(h));
        out[c * K3 + r] = h;
        out[c * K3 + K + r] = l;
        out[c * K3 + 2 * K + r] = h;
    }
}

static void gemm_minus_split16(bemhip_ctx* ctx, int m, int n, int k, const float* A, int lda, const float* B,
                               int ldb, float* C, int ldc) {
    __half* acat = static_cast<__half*>(ctx->split_a.ensure(sizeof(__half) * 3 * (size_t)m * k));
    __half* bcat = static_cast<__half*>(ctx->split_b.ensure(sizeof(__half) * 3 * (size_t)k * n));
    unsigned int* mx = static_cast<unsigned int*>(ctx->split_mx.ensure(64));
    float* alpha = reinterpret_cast<float*>(mx + 4);
    hipStream_t st = ctx->stream;
    k_fill_u32<<<1, 64, 0, st>>>(mx, 2, 0u);
    const int gb = 1024;
    k_block_absmax<<<gb, 256, 0, st>>>(m, k, A, lda, mx);
    k_block_absmax<<<gb, 256, 0, st>>>(k, n, B, ldb, mx + 1);
    k_split_alpha<<<1, 1, 0, st>>>(mx, alpha);
    k_split_a<<<gb, 256, 0, st>>>(m, k, A, lda, mx, acat);
    k_split_b<<<gb, 256, 0, st>>>(k, n, B, ldb, mx, bcat);
    BH_HIP(hipGetLastError());
    static const float one = 1.0f;
    BH_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_device));
    float* d_one = alpha + 1;
    k_fill_u32<<<1, 1, 0, st>>>(reinterpret_cast<unsigned int*>(d_one), 1, __builtin_bit_cast(unsigned int, one));
    rocblas_status s = rocblas_gemm_ex(ctx->blas, rocblas_operation_none, rocblas_operation_none, m, n, 3 * k, alpha,
                                       acat, rocblas_datatype_f16_r, m, bcat, rocblas_datatype_f16_r, 3 * k, d_one, C,
                                       rocblas_datatype_f32_r, ldc, C, rocblas_datatype_f32_r, ldc,
                                       rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
    BH_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host));
    BH_BLAS(s);
}

// timed trailing update
template <class T>
static void trailing_update(bemhip_ctx* ctx, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C,
                            int ldc, bool split16) {
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
        BH_HIP(hipEventRecord(e0, ctx->stream));
    }
    if constexpr (std::is_same<T, float>::value) {
        if (split16) gemm_minus_split16(ctx, m, n, k, A, lda, B, ldb, C, ldc);
        else gemm_minus(ctx->blas, m, n, k, A, lda, B, ldb, C, ldc);
    } else {
        gemm_minus(ctx->blas, m, n, k, A, lda, B, ldb, C, ldc);
    }
    if (ctx->timing) BH_HIP(hipEventRecord(e1, ctx->stream));
}

// sums the recorded trailing-update timings (after the stream has drained)
static void collect_gemm_timing(bemhip_ctx* ctx) {
    for (size_t i = 0; i < ctx->gemm_ev_used; ++i) {
        float ms = 0.0f;
        BH_HIP(hipEventElapsedTime(&ms, ctx->gemm_ev[2 * i], ctx->gemm_ev[2 * i + 1]));
        ctx->gemm_ms += ms;
        ctx->gemm_flops += ctx->gemm_fl[i];
        ctx->gemm_launches += 1;
    }
    ctx->gemm_ev_used = 0;
}

// ------------------------------------------------------------ factorise --
// A stream with the rocBLAS handle bound to it.
struct Lane {
    hipStream_t s;
    rocblas_handle h;
};

template <class T>
static void leaf(bemhip_ctx* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base) {
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
static void panel_rec(bemhip_ctx* ctx, Lane L, int m, int n, T* A, int lda, int* ipiv, int roff, unsigned int& base) {
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

// Blocked right-looking GEPP with one block of look-ahead: while the main
// stream applies step k's update to the trailing matrix, the panel of step
// k+1 (whose columns were updated first) is factored on the second stream.
template <class T> int lu_factor(bemhip_ctx* ctx, int N, T* A, int* ipiv, bool split16) {
    ctx->lu.panel_ws.ensure(PanelWs::BYTES);
    char* ws = ctx->lu.panel_ws.as<char>();
    const Lane main{ctx->stream, ctx->blas};
    const Lane side = ctx->lookahead ? Lane{ctx->stream2, ctx->blas2} : main;
    hipEvent_t ev_ready = ctx->ev_la[0], ev_panel = ctx->ev_la[1];
    k_fill_u32<<<1, 64, 0, main.s>>>(reinterpret_cast<unsigned int*>(ws), 16, 0u);
    BH_HIP(hipGetLastError());
    unsigned int base = 0;
    const int lda = N;
    // panel 0
    BH_HIP(hipEventRecord(ev_ready, main.s));
    BH_HIP(hipStreamWaitEvent(side.s, ev_ready, 0));
    panel_rec(ctx, side, N, std::min(LU_NB, N), A, lda, ipiv, 0, base);
    BH_HIP(hipEventRecord(ev_panel, side.s));
    for (int k = 0; k < N; k += LU_NB) {
        const int jb = std::min(LU_NB, N - k);
        T* Akk = A + (size_t)k * lda + k;
        BH_HIP(hipStreamWaitEvent(main.s, ev_panel, 0));   // panel k factored
        laswp(main, k, A + k, lda, ipiv + k, 0, jb, k);
        const int rest = N - k - jb;
        if (rest <= 0) break;
        T* A12 = A + (size_t)(k + jb) * lda + k;
        laswp(main, rest, A12, lda, ipiv + k, 0, jb, k);
        trsm_llnu(main.h, jb, rest, Akk, lda, A12, lda);
        const int jb2 = std::min(LU_NB, rest);
        // next panel's columns first, then hand them to the side stream
        trailing_update<T>(ctx, rest, jb2, jb, Akk + jb, lda, A12, lda, A12 + jb, lda, split16);
        BH_HIP(hipEventRecord(ev_ready, main.s));
        BH_HIP(hipStreamWaitEvent(side.s, ev_ready, 0));
        panel_rec(ctx, side, rest, jb2, A12 + jb, lda, ipiv + k + jb, k + jb, base);
        BH_HIP(hipEventRecord(ev_panel, side.s));
        if (rest > jb2)
            trailing_update<T>(ctx, rest, rest - jb2, jb, Akk + jb, lda, A12 + (size_t)jb2 * lda, lda,
                               A12 + (size_t)jb2 * lda + jb, lda, split16);
    }
    BH_HIP(hipStreamWaitEvent(main.s, ev_panel, 0));
    unsigned int flags[3] = {0, 0, 0};
    BH_HIP(hipMemcpyAsync(flags, ws, 12, hipMemcpyDeviceToHost, main.s));
    BH_HIP(hipStreamSynchronize(main.s));
    collect_gemm_timing(ctx);
    BH_CHECK(!(flags[1] & 1u), BEMHIP_ERR_HIP, "panel grid barrier timed out");
    if (flags[1] & 2u) return (int)flags[2];  // 1-based first zero pivot
    return 0;
}

template int lu_factor<double>(bemhip_ctx*, int, double*, int*, bool);
template int lu_factor<float>(bemhip_ctx*, int, float*, int*, bool);

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
    for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k]]);
}

template <class T> void lu_solve(bemhip_ctx* ctx, int N, const T* LU, const int* perm, double* x) {
    hipStream_t st = ctx->stream;
    double* tmp = static_cast<double*>(ctx->vec[7].ensure(sizeof(double) * N));
    BH_HIP(hipMemcpyAsync(tmp, x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    k_gather<<<(N + 255) / 256, 256, 0, st>>>(N, perm, tmp, x);
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
template void lu_solve<double>(bemhip_ctx*, int, const double*, const int*, double*);
template void lu_solve<float>(bemhip_ctx*, int, const float*, const int*, double*);

// ------------------------------------------------------ gemv, reductions --
constexpr int GV_CHUNKS = 32;

__global__ __launch_bounds__(256) void k_gemv_part(int N, const double* __restrict__ A, const double* __restrict__ x,
                                                   double* __restrict__ part, int chunk) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = blockIdx.y * chunk, c1 = min(N, c0 + chunk);
    if (r >= N) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = c0;
    const double* col = A + (size_t)c0 * N + r;
    for (; c + 4 <= c1; c += 4, col += 4 * (size_t)N) {
        a0 += col[0] * x[c];
        a1 += col[(size_t)N] * x[c + 1];
        a2 += col[2 * (size_t)N] * x[c + 2];
        a3 += col[3 * (size_t)N] * x[c + 3];
    }
    for (; c < c1; ++c, col += N) a0 += col[0] * x[c];
    part[(size_t)blockIdx.y * N + r] = (a0 + a1) + (a2 + a3);
}

__global__ __launch_bounds__(256) void k_gemv_finish(int N, int chunks, const double* __restrict__ part,
                                                     const double* __restrict__ b, double* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int k = 0; k < chunks; ++k) s += part[(size_t)k * N + i];
    r[i] = b ? b[i] - s : s;
}

void gemv64(bemhip_ctx* ctx, int N, const double* A, const double* x, const double* b, double* r) {
    const int chunks = std::min(GV_CHUNKS, N);
    const int chunk = (N + chunks - 1) / chunks;
    const int nch = (N + chunk - 1) / chunk;
    double* part = static_cast<double*>(ctx->gemv_part.ensure(sizeof(double) * (size_t)nch * N));
    k_gemv_part<<<dim3((N + 255) / 256, nch), 256, 0, ctx->stream>>>(N, A, x, part, chunk);
    k_gemv_finish<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, nch, part, b, r);
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
static double reduce(bemhip_ctx* ctx, int n, const double* x, const double* y) {
    const int nb = 256;
    double* part = static_cast<double*>(ctx->scal.ensure(sizeof(double) * 1024));
    k_reduce_part<OP><<<nb, 256, 0, ctx->stream>>>(n, x, y, part);
    k_reduce_part<OP == 0 ? 0 : 2><<<1, 256, 0, ctx->stream>>>(nb, part, nullptr, part + 512);
    BH_HIP(hipMemcpyAsync(ctx->host_scal, part + 512, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->host_scal[0];
}
double dev_absmax(bemhip_ctx* ctx, int n, const double* x) { return reduce<0>(ctx, n, x, nullptr); }
double dev_dot(bemhip_ctx* ctx, int n, const double* x, const double* y) { return reduce<1>(ctx, n, x, y); }

// y = a*x + b*y (+ c*z when z)
__global__ __launch_bounds__(256) void k_axpby(int n, double a, const double* __restrict__ x, double b,
                                               double* __restrict__ y, double c, const double* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = a * x[i] + b * y[i];
    if (z) v += c * z[i];
    y[i] = v;
}
static void axpby(bemhip_ctx* ctx, int n, double a, const double* x, double b, double* y, double c = 0.0,
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

// Eigen 3.4 bicgstab() with DiagonalPreconditioner, zero initial guess
static void bicgstab(bemhip_ctx* ctx, int N, const double* A, const double* rhs, double* x, double tol, int maxit,
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
static int* make_perm(bemhip_ctx* ctx, int N, const int* d_ipiv) {
    ctx->h_ipiv.resize(N);
    BH_HIP(hipMemcpyAsync(ctx->h_ipiv.data(), d_ipiv, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    pivots_to_perm(ctx->h_ipiv, ctx->h_perm);
    int* d = static_cast<int*>(ctx->perm.ensure(sizeof(int) * N));
    BH_HIP(hipMemcpyAsync(d, ctx->h_perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return d;
}

static float elapsed(bemhip_ctx* ctx, hipEvent_t a, hipEvent_t b) {
    BH_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void solve_device(bemhip_ctx* ctx, int N, double* G64, float* G32, const double* b, double* x, int solver,
                  double tol, int maxit, bemhip_solve_stats* st) {
    hipStream_t sm = ctx->stream;
    int* ipiv = static_cast<int*>(ctx->lu.ipiv.ensure(sizeof(int) * N));
    ctx->gemm_ms = 0.0; ctx->gemm_flops = 0.0; ctx->gemm_launches = 0;
    hipEvent_t e0, e1, e2;
    BH_HIP(hipEventCreate(&e0)); BH_HIP(hipEventCreate(&e1)); BH_HIP(hipEventCreate(&e2));
    BH_HIP(hipEventRecord(e0, sm));
    st->iterations = 0; st->converged = 1; st->correction = 0.0; st->residual = 0.0;
    if (solver == BEMHIP_SOLVE_BICGSTAB64) {
        BH_HIP(hipEventRecord(e1, sm));
        bicgstab(ctx, N, G64, b, x, tol, maxit, st);
    } else if (solver == BEMHIP_SOLVE_LU64) {
        const int info = lu_factor<double>(ctx, N, G64, ipiv, false);
        BH_CHECK(info == 0, BEMHIP_ERR_SINGULAR, "zero pivot in column " + std::to_string(info));
        int* perm = make_perm(ctx, N, ipiv);
        BH_HIP(hipEventRecord(e1, sm));
        BH_HIP(hipMemcpyAsync(x, b, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        lu_solve<double>(ctx, N, G64, perm, x);
    } else if (solver == BEMHIP_SOLVE_LU32_IR64 || solver == BEMHIP_SOLVE_LU16X3_IR64) {
        BH_CHECK(G32 != nullptr, BEMHIP_ERR_ARG, "mixed solve needs the FP32 copy of G");
        const int info = lu_factor<float>(ctx, N, G32, ipiv, solver == BEMHIP_SOLVE_LU16X3_IR64);
        BH_CHECK(info == 0, BEMHIP_ERR_SINGULAR, "zero pivot in column " + std::to_string(info));
        int* perm = make_perm(ctx, N, ipiv);
        BH_HIP(hipEventRecord(e1, sm));
        if (tol <= 0.0) tol = 1e-12;
        if (maxit <= 0) maxit = 30;
        BH_HIP(hipMemcpyAsync(x, b, sizeof(double) * N, hipMemcpyDeviceToDevice, sm));
        lu_solve<float>(ctx, N, G32, perm, x);
        double* d = static_cast<double*>(ctx->vec[6].ensure(sizeof(double) * N));
        int it = 0;
        bool conv = false;
        double cn = 0.0;
        while (it < maxit) {
            gemv64(ctx, N, G64, x, b, d);   // d = b - G x  (FP64 residual)
            lu_solve<float>(ctx, N, G32, perm, d);
            axpby(ctx, N, 1.0, d, 1.0, x);
            ++it;
            const double dn = dev_absmax(ctx, N, d), xn = dev_absmax(ctx, N, x);
            cn = xn > 0.0 ? dn / xn : dn;
            if (!std::isfinite(cn)) break;
            if (cn <= tol) { conv = true; break; }
        }
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
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
    // final FP64 residual |b - G x| / |b| (not for LU64, whose G was overwritten)
    if (solver != BEMHIP_SOLVE_LU64) {
        double* r = static_cast<double*>(ctx->vec[5].ensure(sizeof(double) * N));
        gemv64(ctx, N, G64, x, b, r);
        const double rn = dev_dot(ctx, N, r, r), bn = dev_dot(ctx, N, b, b);
        st->residual = bn > 0.0 ? std::sqrt(rn / bn) : std::sqrt(rn);
    }
}

}  // namespace bemhip
