// Triangular inverse of an LU panel's diagonal block (the unit-lower L11 or
// the non-unit upper U11 of an n x n block stored as one column-major square),
// for the blocked LU's products U12 = L11^-1 A12 and L21 = A21 U11^-1.
//
// The reference's LU (Eigen PartialPivLU, Simulation.cpp:57-61) solves these
// with triangular solves; the MI355X LU multiplies by the inverse instead, so
// both products become matrix-core GEMMs.  The inverse is built in two kinds
// of launch on the caller's stream:
//   1. k_trinv_diag: every 32 x 32 diagonal block inverted in registers, one
//      half-wave per block (lane i holds row i of the block's inverse), and
//      the zero triangle of W written;
//   2. per doubling level s = 32, 64, ... < n, two batched GEMMs that merge
//      neighbouring inverted groups of size s (skipping the K range where
//      the triangular operand is zero):
//        lower:  [X_AA 0; C X_BB]  ->  X_BA = -X_BB (C X_AA)
//        upper:  [X_AA C; 0 X_BB]  ->  X_AB = -X_AA (C X_BB)
//      C being the block of the panel between the two groups.
// Both triangles of one panel are built by the same 11 launches (n = 1024).
// Flops: n^3 / 3 a triangle, about n s^2 at level s; the launches are latency
// bound (a few memory round trips each), not bandwidth or compute bound.
#include "internal.hpp"

namespace bemhip {

constexpr int TI_B = 32;   // diagonal block of the first launch

// W (ld n): the inverse's 32 x 32 diagonal blocks and the zero triangle of
// W.  Each half-wave inverts one block (eight a work-group), and TI_ZWG more
// work-groups write the zero triangle:
// lane i holds row i of the block's inverse, and step k broadcasts the
// finished row k through LDS to the rows it updates (right-looking
// substitution; one barrier a step, the broadcast rows double-buffered).
//   lower, unit:  X[i,:] = e_i - sum_{k<i} L[i][k] X[k,:]
//   upper:        X[i,:] = (e_i - sum_{k>i} U[i][k] X[k,:]) / U[i][i]
// grid.y selects the triangle: 0 the lower into Wl, 1 the upper into Wu (a
// null W skips that triangle).
constexpr int TI_ZWG = 16;   // work-groups a triangle that write its zero triangle

// grid.z: problem of a batch (A and W sA, sW elements further each)
template <class T>
__global__ __launch_bounds__(256) void k_trinv_diag(int n, const T* __restrict__ A, int lda, T* __restrict__ Wl,
                                                    T* __restrict__ Wu, long long sA, long long sW) {
    __shared__ T s[8][TI_B][TI_B + 1];   // row-major copy of each block: s[b][i][k] = A(i0 + i, i0 + k)
    __shared__ T rb[2][8][TI_B];         // broadcast row of each block, by step parity
    const bool lower = blockIdx.y == 0;
    T* __restrict__ W = lower ? Wl : Wu;
    if (!W) return;   // the same for the whole work-group
    A += blockIdx.z * sA;
    W += blockIdx.z * sW;
    const int nbk = (n + TI_B - 1) / TI_B;
    const int ndiag = (nbk + 7) / 8;
    if ((int)blockIdx.x >= ndiag) {
        // the zero triangle outside the diagonal blocks: rows above column c's
        // diagonal block (lower) or below it (upper); columns dealt
        // round-robin so every work-group writes about the same amount
        for (int c = (int)blockIdx.x - ndiag; c < n; c += TI_ZWG) {
            const int cb = c / TI_B * TI_B;
            const int r0 = lower ? 0 : min(n, cb + TI_B), r1 = lower ? cb : n;
            T* col = W + (size_t)c * n;
            for (int r = r0 + (int)threadIdx.x; r < r1; r += 256) col[r] = T(0);
        }
        return;
    }
    const int lane = threadIdx.x & 63, hb = threadIdx.x >> 5;   // hb: this half-wave's block slot
    const int bi = blockIdx.x * 8 + hb;
    const int i0 = bi * TI_B;
    const int bs = bi < nbk ? min(TI_B, n - i0) : 0;
    const int i = lane & (TI_B - 1);
    {
        // all 32 loads in flight before the first LDS store
        T v[TI_B];
#pragma unroll
        for (int c = 0; c < TI_B; ++c) v[c] = (i < bs && c < bs) ? A[(size_t)(i0 + c) * lda + i0 + i] : T(0);
#pragma unroll
        for (int c = 0; c < TI_B; ++c) s[hb][i][c] = v[c];
    }
    // rows and columns beyond bs: a zero block with a unit diagonal, so they
    // stay the identity's
    T x[TI_B];
#pragma unroll
    for (int c = 0; c < TI_B; ++c) x[c] = c == i ? T(1) : T(0);
    __syncthreads();
    // the step loop stays rolled: unrolled, its 32 copies overflow the
    // instruction cache
#pragma unroll 1
    for (int q = 0; q < TI_B; ++q) {
        const int k = lower ? q : TI_B - 1 - q;
        if (i == k) {
            if (!lower) {
                const T r = T(1) / (k < bs ? s[hb][k][k] : T(1));
#pragma unroll
                for (int c = 0; c < TI_B; ++c) x[c] *= r;
            }
#pragma unroll
            for (int c = 0; c < TI_B; ++c) rb[q & 1][hb][c] = x[c];
        }
        __syncthreads();
        if (lower ? i > k : i < k) {
            const T f = s[hb][i][k];
#pragma unroll
            for (int c = 0; c < TI_B; ++c) x[c] -= f * rb[q & 1][hb][c];
        }
    }
    if (bs == 0) return;
    for (int c = 0; c < bs; ++c)
        if (i < bs) W[(size_t)(i0 + c) * n + i0 + i] = x[c];
}

// which K range of a merge product can be non-zero for the 64 x 64 output
// tile at (m0, n0): the triangular factor's zero half is skipped
enum TiSkip : int {
    TI_FULL = 0,
    TI_B_LOWER = 1,   // B lower triangular: k >= n0
    TI_A_LOWER = 2,   // A lower triangular: k < m0 + 64
    TI_B_UPPER = 3,   // B upper triangular: k < n0 + 64
    TI_A_UPPER = 4,   // A upper triangular: k >= m0
};

// one batched product C_z = alpha A_z B_z (column-major, A_z m x k, B_z k x
// n) over `pairs` pairs; pair z's operands start at p + z * stride, and the
// last pair has its own (ragged) m, n, k
template <class T> struct TiTask {
    const T* a;
    const T* b;
    T* c;
    int lda, ldb, ldc;
    int64_t sa, sb, sc;
    int M, N, K, Ml, Nl, Kl, pairs, skip;
    T alpha;
    // a batch of problems: the pairs of problem q start qa, qb, qc elements further
    int64_t qa, qb, qc;
};

constexpr int TI_KS = 32;   // K slice of the merge products

// grid.z runs over the pairs of task 0 and then those of task 1 (the lower
// and the upper triangle of the same doubling level).  64 x 64 tile per
// work-group, 4 x 4 outputs a thread, 32-deep slices of A and B staged
// through LDS with the next slice's loads in flight during the current
// slice's products.  At the largest level (s = n / 2) the products are
// bound by the vector FP32 rate of the compute units the concurrent trailing
// update leaves free.
// grid.z: nbat problems x (task 0's pairs, then task 1's)
template <class T>
__global__ __launch_bounds__(256) void k_trinv_gemm(TiTask<T> t0, TiTask<T> t1) {
    const int zz = (int)blockIdx.z % (t0.pairs + t1.pairs), q = (int)blockIdx.z / (t0.pairs + t1.pairs);
    const bool second = zz >= t0.pairs;
    const TiTask<T>& t = second ? t1 : t0;
    const int z = second ? zz - t0.pairs : zz;
    const bool last = z == t.pairs - 1;
    const int m = last ? t.Ml : t.M, n = last ? t.Nl : t.N, k = last ? t.Kl : t.K;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    if (m0 >= m || n0 >= n) return;   // the same for the whole work-group
    const T* __restrict__ A = t.a + (int64_t)z * t.sa + q * t.qa;
    const T* __restrict__ B = t.b + (int64_t)z * t.sb + q * t.qb;
    T* __restrict__ C = t.c + (int64_t)z * t.sc + q * t.qc;
    int kb = 0, ke = k;
    if (t.skip == TI_B_LOWER) kb = n0;
    else if (t.skip == TI_A_UPPER) kb = m0;
    else if (t.skip == TI_A_LOWER) ke = min(k, m0 + 64);
    else if (t.skip == TI_B_UPPER) ke = min(k, n0 + 64);
    __shared__ __align__(16) T As[TI_KS][64];
    __shared__ T Bs[TI_KS][65];
    // thread (tx, ty): rows m0 + 4 tx .. + 3 (one 16-byte LDS read of A a k),
    // columns n0 + ty + 16 j
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    // this thread's loads of a slice: A as (kk, mm) pairs, B as (kk, nn) pairs
    T ra[8], rbv[8];
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = tid + 256 * r;
            const int mm = e & 63, kk = e >> 6;
            ra[r] = (m0 + mm < m && k0 + kk < ke) ? A[(size_t)(k0 + kk) * t.lda + m0 + mm] : T(0);
            const int kq = e & (TI_KS - 1), nn = e >> 5;
            rbv[r] = (k0 + kq < ke && n0 + nn < n) ? B[(size_t)(n0 + nn) * t.ldb + k0 + kq] : T(0);
        }
    };
    T acc[4][4] = {};
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += TI_KS) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = tid + 256 * r;
            As[e >> 6][e & 63] = ra[r];
            Bs[e & (TI_KS - 1)][e >> 5] = rbv[r];
        }
        __syncthreads();
        if (k0 + TI_KS < ke) load(k0 + TI_KS);
#pragma unroll
        for (int kk = 0; kk < TI_KS; ++kk) {
            T av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][4 * tx + i];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) bv[jj] = Bs[kk][ty + 16 * jj];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) acc[i][jj] += av[i] * bv[jj];
        }
        __syncthreads();
    }
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int col = n0 + ty + 16 * jj;
        if (col >= n) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + 4 * tx + i;
            if (row < m) C[(size_t)col * t.ldc + row] = t.alpha * acc[i][jj];
        }
    }
}

size_t tri_inv_tmp_elems(int n) {
    size_t mx = 0;
    for (int s = TI_B; s < n; s *= 2) {
        const size_t pairs = (size_t)(n - s + 2 * s - 1) / (2 * s);
        mx = std::max(mx, pairs * (size_t)s * (size_t)s);
    }
    return std::max<size_t>(mx, 1);
}

template <class T>
void tri_inv2(hipStream_t st, int n, const T* A, int lda, T* Wl, T* Wu, T* tmp_l, T* tmp_u, const TriBat* bat) {
    if (n <= 0 || (!Wl && !Wu)) return;
    const TriBat one{1, 0, 0, 0};
    if (!bat) bat = &one;
    BH_CHECK(bat->n >= 1 && bat->n <= 1024, BEMHIP_ERR_ARG, "tri_inv2: bad batch");
    const int64_t qA = bat->sA, qW = bat->sW, qT = bat->sT;
    const int nbk = (n + TI_B - 1) / TI_B;
    k_trinv_diag<T><<<dim3((nbk + 7) / 8 + TI_ZWG, 2, bat->n), 256, 0, st>>>(n, A, lda, Wl, Wu, qA, qW);
    BH_HIP(hipGetLastError());
    const TiTask<T> none{};
    for (int s = TI_B; s < n; s *= 2) {
        // pairs of neighbouring groups [2ts, 2ts + s) and [2ts + s, 2ts + 2s)
        // that have a second group; the last one's may be short (sb)
        const int pairs = (n - s + 2 * s - 1) / (2 * s);
        const int sb = std::min(s, n - (2 * (pairs - 1) * s + s));
        const int64_t sa = 2 * (int64_t)s * (lda + 1), sw = 2 * (int64_t)s * (n + 1), st2 = (int64_t)s * s;
        // lower: tmp = C X_AA (sb x s) with C = A(b0.., a0..), then
        // X_BA = -X_BB tmp with X_BB = W(b0.., b0..)
        const TiTask<T> l1{A + s, Wl, tmp_l, lda, n, s, sa, sw, st2, s, s, s, sb, s, s, pairs, TI_B_LOWER, T(1),
                           qA, qW, qT};
        const TiTask<T> l2{Wl + (size_t)s * n + s, tmp_l, Wl + s, n, s, n, sw, st2, sw,
                           s, s, s, sb, s, sb, pairs, TI_A_LOWER, T(-1), qW, qT, qW};
        // upper: tmp = C X_BB (s x sb) with C = A(a0.., b0..), then
        // X_AB = -X_AA tmp
        const TiTask<T> u1{A + (size_t)s * lda, Wu + (size_t)s * n + s, tmp_u, lda, n, s, sa, sw, st2,
                           s, s, s, s, sb, sb, pairs, TI_B_UPPER, T(1), qA, qW, qT};
        const TiTask<T> u2{Wu, tmp_u, Wu + (size_t)s * n, n, s, n, sw, st2, sw,
                           s, s, s, s, sb, s, pairs, TI_A_UPPER, T(-1), qW, qT, qW};
        const int z = ((Wl ? pairs : 0) + (Wu ? pairs : 0)) * bat->n;
        const dim3 grid((s + 63) / 64, (s + 63) / 64, z);
        k_trinv_gemm<T><<<grid, 256, 0, st>>>(Wl ? l1 : u1, Wl && Wu ? u1 : none);
        k_trinv_gemm<T><<<grid, 256, 0, st>>>(Wl ? l2 : u2, Wl && Wu ? u2 : none);
        BH_HIP(hipGetLastError());
    }
}
template void tri_inv2<float>(hipStream_t, int, const float*, int, float*, float*, float*, float*, const TriBat*);
template void tri_inv2<double>(hipStream_t, int, const double*, int, double*, double*, double*, double*,
                               const TriBat*);

}  // namespace bemhip
