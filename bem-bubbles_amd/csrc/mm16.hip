// fp16 matrix-core products of the mixed-precision LU on gfx950
// (v_mfma_f32_32x32x16_f16: fp16 operands, FP32 accumulation).
//
// Replaces, for the fp16 LU of Simulation::solve_system (Simulation.cpp:57-61,
// Eigen::PartialPivLU), every GEMM-shaped product of the blocked factorisation:
// the trailing update A22 -= L21 U12, the block row U12 = L11^-1 A12 and the
// rows below a panel L21 = A21 U11^-1.
//
// Operands live in HBM as fp16 "fragment tiles": the matrix is cut into blocks
// of 32 rows (or columns) x 16 k, and each block is stored as the 64 lanes x
// 8 halves (16 B per lane) that one MFMA operand register quad holds: lane l
// carries row (l & 31) and k = 8 (l >> 5) .. +7.  A block is 1 KB contiguous,
// so the GEMM moves it HBM -> LDS with one direct-to-LDS dwordx4 load per wave
// (no VGPR staging) and LDS -> registers with one conflict-free ds_read_b128
// per lane.  Blocks are ordered row-block major: block (rb, kb) at
// (rb * KB + kb) KB, KB = padded K / 16.  The pack kernels write these tiles
// from FP32 column-major matrices (scaled by a power of two, zero padded to
// 256 rows and 64 k) and log max |scaled value| so that the host can detect an
// fp16 overflow after the factorisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "internal.hpp"

namespace bemhip {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MM_BKB = MM_BK / 16;          // 16-k blocks per k-step (4)

// One direct-to-LDS load: each lane's 16 bytes at g land at the wave-uniform
// LDS address s + 16 lane.  Written as inline asm (not the builtin) so that
// the compiler does not order later LDS reads of the other stage behind it;
// the kernel waits for it explicitly (vmcnt) before a barrier.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ inline void lds_dma16(const h8* g, h8* s) {
    const unsigned m = (unsigned)(size_t)(__attribute__((address_space(3))) h8*)s;
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(m), "v"(g) : "memory", "m0");
}
#pragma clang diagnostic pop
constexpr int MM_RB = MM_TILE / 32;         // 32-row blocks per macro tile (8)

// ------------------------------------------------------------------ pack --
// Both pack kernels go through LDS so that the FP32 source is read along its
// contiguous dimension and every wave writes whole 1 KB fragment blocks.
// maxlog: max |scale x| merged as float bits (NaN and Inf log as +Inf);
// amax64: max |x| (unscaled) merged as the bits of a non-negative double (NaN
// orders above every number).
__device__ inline void pack_logs(float m, double am, unsigned int* maxlog, unsigned long long* amax64) {
    if (maxlog) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(maxlog, __float_as_uint(m));
    }
    if (amax64) {
        unsigned long long b = (unsigned long long)__double_as_longlong(am);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(b, off);
            b = o > b ? o : b;
        }
        if ((threadIdx.x & 63) == 0 && b > 0ull) atomicMax(amax64, b);
    }
}

// per-problem offsets of a batched pack (blockIdx.z = problem): source and
// tiles strides (floats, h8), log stride, scale exponents 2^(ek e[b])
struct PackBat {
    long long sx, so;
    int ls, ek;
    const int* e;
};

// A operand: X is R x K column-major (element (r, k) = X[r + k ld]).  A
// work-group packs 256 rows x 16 k (8 fragment blocks): grid (RP / 256, KB).
__global__ __launch_bounds__(256) void k_pack_a16(int R, int K, int KB, const float* __restrict__ X, long ld,
                                                  float scale, h8* __restrict__ out, unsigned int* __restrict__ maxlog,
                                                  unsigned long long* __restrict__ amax64, PackBat pb) {
    __shared__ float s[16][256];
    const int r0 = blockIdx.x * 256, kb = blockIdx.y, k0 = kb * 16, t = threadIdx.x;
    {
        const int b = blockIdx.z;   // problem of a batch
        X += b * pb.sx;
        out += b * pb.so;
        if (maxlog) maxlog += b * pb.ls;
        if (amax64) amax64 += b * pb.ls;
        if (pb.e) scale = ldexpf(scale, pb.ek * pb.e[b]);
    }
    const int r = r0 + t;
#pragma unroll
    for (int k = 0; k < 16; ++k) s[k][t] = (r < R && k0 + k < K) ? X[(size_t)(k0 + k) * ld + r] : 0.0f;
    __syncthreads();
    const int w = t >> 6, l = t & 63;
    float m = 0.0f;
    double am = 0.0;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int bl = 2 * w + u;   // local row block
        const int rl = bl * 32 + (l & 31), kk = (l >> 5) * 8;
        h8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float x = s[kk + e][rl];
            const float y = x * scale;
            const float a = fabsf(y);
            m = (a <= 3.0e38f) ? fmaxf(m, a) : __builtin_inff();
            am = fmax(am, fabs((double)x));
            if (x != x) am = __longlong_as_double(0x7ff8000000000000ll);
            v[e] = (_Float16)y;
        }
        out[((size_t)(blockIdx.x * MM_RB + bl) * KB + kb) * 64 + l] = v;
    }
    pack_logs(m, am, maxlog, amax64);
}

// B operand: X is K x C column-major (element (k, c) = X[k + c ld]).  A
// work-group packs 32 columns x 128 k (8 fragment blocks): grid (KP / 128,
// CP / 32), the k chunks of a column block in consecutive work-groups (one
// DRAM run per column).  copy (optional, ld ldc) receives the unscaled values.
constexpr int PB_K = 128;
__global__ __launch_bounds__(256) void k_pack_b16(int C, int K, int KB, const float* __restrict__ X, long ld,
                                                  float scale, h8* __restrict__ out, unsigned int* __restrict__ maxlog,
                                                  float* __restrict__ copy, long ldc, PackBat pb) {
    __shared__ float s[32][PB_K + 4];
    const int c0 = blockIdx.y * 32, k0 = blockIdx.x * PB_K, t = threadIdx.x;
    {
        const int b = blockIdx.z;   // problem of a batch (no copy in a batch)
        X += b * pb.sx;
        out += b * pb.so;
        if (maxlog) maxlog += b * pb.ls;
        if (pb.e) scale = ldexpf(scale, pb.ek * pb.e[b]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int idx = t + 256 * j, c = idx >> 5, q = idx & 31;
        const int col = c0 + c, k = k0 + 4 * q;
        float4 v = {0.f, 0.f, 0.f, 0.f};
        if (col < C) {
            const float* src = X + (size_t)col * ld + k;
            if (k + 3 < K) {
                __builtin_memcpy(&v, src, 16);
                if (copy) __builtin_memcpy(copy + (size_t)col * ldc + k, &v, 16);
            } else {
                float e4[4] = {0.f, 0.f, 0.f, 0.f};
                for (int e = 0; e < 4; ++e)
                    if (k + e < K) {
                        e4[e] = src[e];
                        if (copy) copy[(size_t)col * ldc + k + e] = e4[e];
                    }
                v = float4{e4[0], e4[1], e4[2], e4[3]};
            }
        }
        *reinterpret_cast<float4*>(&s[c][4 * q]) = v;
    }
    __syncthreads();
    const int w = t >> 6, l = t & 63;
    float m = 0.0f;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int kl = (2 * w + u) * 16 + (l >> 5) * 8;   // local k of this lane's 8
        const float4 p0 = *reinterpret_cast<const float4*>(&s[l & 31][kl]);
        const float4 p1 = *reinterpret_cast<const float4*>(&s[l & 31][kl + 4]);
        const float x[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
        h8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float y = x[e] * scale;
            const float a = fabsf(y);
            m = (a <= 3.0e38f) ? fmaxf(m, a) : __builtin_inff();
            v[e] = (_Float16)y;
        }
        const int kb = blockIdx.x * (PB_K / 16) + 2 * w + u;
        if (kb < KB) out[((size_t)blockIdx.y * KB + kb) * 64 + l] = v;
    }
    pack_logs(m, 0.0, maxlog, nullptr);
}

void mm16_pack(hipStream_t st, bool kcontig, int R, int K, const float* X, long ld, float scale, void* tiles,
               unsigned int* maxlog, float* copy, long ldc, unsigned long long* amax64, const Mm16Bat* bat) {
    if (R <= 0 || K <= 0) return;
    const int RP = mm_pad(R, MM_TILE), KB = mm_pad(K, MM_BK) / 16;
    PackBat pb{0, 0, 0, 0, nullptr};
    int nb = 1;
    if (bat) {
        BH_CHECK(bat->n >= 1 && bat->n <= 65535 && bat->sa % 16 == 0 && copy == nullptr, BEMHIP_ERR_ARG,
                 "mm16_pack: bad batch");
        pb = PackBat{bat->sx, bat->sa / 16, bat->ls, bat->ek, bat->e};
        nb = bat->n;
    }
    if (kcontig) {
        BH_CHECK(amax64 == nullptr, BEMHIP_ERR_ARG, "mm16_pack: no max |x| for B operands");
        k_pack_b16<<<dim3((KB * 16 + PB_K - 1) / PB_K, RP / 32, nb), 256, 0, st>>>(
            R, K, KB, X, ld, scale, static_cast<h8*>(tiles), maxlog, copy, ldc, pb);
    } else {
        BH_CHECK(copy == nullptr, BEMHIP_ERR_ARG, "mm16_pack: no copy for A operands");
        k_pack_a16<<<dim3(RP / 256, KB, nb), 256, 0, st>>>(R, K, KB, X, ld, scale, static_cast<h8*>(tiles), maxlog,
                                                            amax64, pb);
    }
    BH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ gemm --
// C (M x N, column-major, ld ldc) -= alpha A B   (EPI = MM_SUB), or
// C = alpha A B                                  (EPI = MM_STORE),
// A: fragment tiles of M rows x K (KBa 16-k blocks per row block), B: fragment
// tiles of K x (columns from 32-column block cb0 on; KBb blocks per column
// block).  nks k-steps of 64 (the tiles are zero padded to them).  TRI = 1: A
// is lower triangular (tile row tm only needs k < 256 (tm + 1)); TRI = 2: B is
// upper triangular (tile column tn only needs k < 256 (tn + 1)).
//
// A work-group (4 waves) owns a 256 x TN tile of C; wave (wm, wn) a 128 x TN/2
// quarter as 4 x TN/64 MFMA 32x32 accumulators (TN FP32 per lane).  Per stage
// of 32 k the work-group stages 16 KB of A and TN/16 KB of B in LDS by
// direct-to-LDS loads, four stages deep (TN = 256, one work-group per compute
// unit) or three (TN = 128: two work-groups share a compute unit, two waves a
// SIMD, so that one work-group's waits and epilogue overlap the other's
// MFMAs; about twice the rate per compute unit where the launch does not fill
// the chip, as the LU's narrow products beside the bulk update).  The launch covers
// the tiles in groups of 8 tile rows, and consecutive work-group ids (which
// the hardware deals round-robin to the 8 XCDs) are remapped so that every XCD
// works through a contiguous range of groups: the work-groups sharing an L2
// share A row blocks and B column blocks.
struct MmArgs {
    const h8* At;        // A tiles, KBa 16-k blocks per row block
    const h8* Bt;        // B tiles, KBb blocks per column block, read from column block cb0
    float* C;
    long ldc;
    float alpha;
    int KBa, KBb, cb0, M, N, nks, tiles_m, tiles_n, per_xcd, wg_xcd;
    // the result (stored, or C after the subtraction) also as fp16 tiles
    // (tkind 1: A tiles, k = column of C, C's rows from tr0 on; 2: B tiles,
    // k = row of C, column blocks from tcb0), tKB blocks per row (column)
    // block, scaled by tscale; logs as the packs
    h8* tout;
    int tkind, tKB, tcb0, tr0;
    float tscale;
    unsigned int* maxlog;
    unsigned long long* amax64;
    // a batch of nbat equally shaped products: problem b's A, B and output
    // tiles start b * sAt, sBt, sT h8 further, its C b * sC floats further;
    // alpha and tscale carry 2^(ek e[b]), 2^(ek2 e[b]); logs ls apart
    int nbat;
    long long sAt, sBt, sC, sT;
    const int* ebat;
    int ek, ek2, ls;
    // a second operand segment (At2 null: none): the k-steps from nks1 on
    // (of 64 k) read A2 = At2 (KBa2 blocks per row block) and B2 = Bt2 (KBb2
    // per column block, from column block cb02) from their k = 0, i.e. the
    // product is A B + A2 B2 over the concatenated k range
    const h8* At2;
    const h8* Bt2;
    int KBa2, KBb2, cb02, nks1;
};

// BULK: the same kernel for the LU's bulk update of two panels (persistent,
// two operand segments), instantiated apart so that profiles list it apart
// s_waitcnt vmcnt with a count known only at run time (one of the values the
// stage schedules below produce)
__device__ inline void vm_wait(int n) {
    switch (n) {
    case 40: asm volatile("s_waitcnt vmcnt(40)" ::: "memory"); break;
    case 32: asm volatile("s_waitcnt vmcnt(32)" ::: "memory"); break;
    case 24: asm volatile("s_waitcnt vmcnt(24)" ::: "memory"); break;
    case 16: asm volatile("s_waitcnt vmcnt(16)" ::: "memory"); break;
    case 12: asm volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
}

template <int EPI, int TRI, bool BULK = false, int TN = 256>
__global__ __launch_bounds__(256, TN == 256 ? 1 : 2) void k_mm16(const MmArgs g) {
    static_assert(TN == 256 || TN == 128, "k_mm16: tiles of 256 x 256 or 256 x 128");
    const long ldc = g.ldc;
    const int KBa = g.KBa, KBb = g.KBb, cb0 = g.cb0, M = g.M, N = g.N, nks = g.nks;
    const int tiles_m = g.tiles_m, tiles_n = g.tiles_n, per_xcd = g.per_xcd, wg_xcd = g.wg_xcd;
    // four stages of [A 16 KB | B 16 KB], 32 k each: three stages' loads are
    // in flight while one is computed (two 64-k stages left the next stage's
    // loads a single stage of compute to arrive in, which they often missed);
    // TN = 128: three stages of [A 16 KB | B 8 KB], 72 KB per work-group
    constexpr int SKB = 2;                          // 16-k blocks per stage
    constexpr int NST = TN == 256 ? 4 : 3;          // stages
    constexpr int NJ = TN / 64;                     // accumulator columns per wave
    constexpr int ABL = MM_RB * SKB, BBL = (TN / 32) * SKB;   // 1 KB blocks per stage
    constexpr int STAGE = (ABL + BBL) * 64;         // h8 per stage
    constexpr int LPS = (ABL + BBL) / 4;            // loads per wave and stage
    // the epilogue's first C chunk is loaded behind the last stage (TN = 256;
    // with two work-groups a compute unit the registers are not there)
    constexpr bool CPRE = TN == 256;
    __shared__ h8 lds[NST * STAGE];
    const int xcd = blockIdx.x & 7;
    float tmax = 0.0f;    // max |tscale x| of the fused tiles
    double tamax = 0.0;   // max |x|
    const int tpp = tiles_m * tiles_n;   // tiles per problem
    for (int t = blockIdx.x >> 3; t < per_xcd; t += wg_xcd) {
    int L = xcd * per_xcd + t;
    if (L >= tpp * g.nbat) break;
    const int bq = L / tpp;   // problem of a batch
    L -= bq * tpp;
    const h8* __restrict__ At = g.At + bq * g.sAt;
    const h8* __restrict__ Bt = g.Bt + bq * g.sBt;
    float* __restrict__ C = g.C + bq * g.sC;
    h8* __restrict__ tout = g.tout ? g.tout + bq * g.sT : nullptr;
    const float alpha = g.ebat ? ldexpf(g.alpha, g.ek * g.ebat[bq]) : g.alpha;
    const float tscale = g.ebat ? ldexpf(g.tscale, g.ek2 * g.ebat[bq]) : g.tscale;
    constexpr int GM = 8;
    const int fm = (L / (GM * tiles_n)) * GM;
    const int gs = min(tiles_m - fm, GM);
    const int q = L - fm * tiles_n;
    const int tm = fm + q % gs, tn = q / gs;
    int kend = nks;
    if (TRI == 1) kend = min(nks, (tm + 1) * (MM_TILE / MM_BK));
    if (TRI == 2) kend = min(nks, (tn + 1) * (TN / MM_BK));
    const int nst = kend * (MM_BKB / SKB);          // stages of this tile
    // the wave index as a scalar: every staging address is then a scalar base
    // plus the lane's 16 bytes
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
    const int wm = w & 1, wn = w >> 1;
    const h8* abase = At + (size_t)tm * MM_RB * KBa * 64;
    const h8* bbase = Bt + (size_t)(cb0 + tn * (TN / 32)) * KBb * 64;
    // the second segment's operands and its first stage (stages of 32 k)
    const h8* abase2 = g.At2 ? g.At2 + (size_t)tm * MM_RB * g.KBa2 * 64 : abase;
    const h8* bbase2 = g.At2 ? g.Bt2 + (size_t)(g.cb02 + tn * (TN / 32)) * g.KBb2 * 64 : bbase;
    const int nst1 = g.At2 ? g.nks1 * (MM_BKB / SKB) : (1 << 30);

    // wave w stages blocks 4w .. 4w+3 of A and TN/64 blocks of B (block a:
    // local row block a >> 1, k block a & 1 of the stage; LDS slot a): LPS
    // loads a lane
    auto issue = [&](int st) {
        h8* sA = lds + (st % NST) * STAGE;
        h8* sB = sA + ABL * 64;
        const bool s2 = st >= nst1;
        const h8* ab = s2 ? abase2 : abase;
        const h8* bb = s2 ? bbase2 : bbase;
        const int kba = s2 ? g.KBa2 : KBa, kbb = s2 ? g.KBb2 : KBb, sk = (s2 ? st - nst1 : st) * SKB;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int a = w * 4 + u;
            const h8* src = ab + ((size_t)(a >> 1) * kba + sk + (a & 1)) * 64;
            lds_dma16(src + l, sA + a * 64);
        }
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            const int b = w * NJ + u;
            const h8* src = bb + ((size_t)(b >> 1) * kbb + sk + (b & 1)) * 64;
            lds_dma16(src + l, sB + b * 64);
        }
    };
    // the epilogue's view of this wave's 128 x 128 quarter of C
    const int row0 = tm * MM_TILE + wm * 128, col0 = tn * TN + wn * (TN / 2);
    const int rr = 4 * (l & 31);   // this lane's 4 rows (of 128) when reading back
    const bool rows_ok = row0 + rr + 3 < M;
    // C of a 32-column chunk is loaded while the one before is transposed and
    // written: 16 loads of 16 bytes per lane in flight
    constexpr int NCV = CPRE ? 2 : 1;   // C chunks in registers
    float4 cv[NCV][16];
    // Lanes whose column or rows fall outside C read a clamped position
    // instead (never used: the epilogue skips such columns and reloads
    // partial row groups), so that the loads need no exec mask or zero fill
    // and the compiler places no wait before them; with fewer than 4 rows
    // the clamped read may touch up to 3 elements past the block's rows.
    const int rclamp = rows_ok ? row0 + rr : max(0, M - 4);
    auto load_chunk = [&](int j, float4* c) {
        if (EPI != MM_SUB) return;
#pragma unroll
        for (int c2 = 0; c2 < 16; ++c2) {
            const int col = min(col0 + j * 32 + 2 * c2 + (l >> 5), N - 1);
            __builtin_memcpy(&c[c2], C + (size_t)col * ldc + rclamp, 16);
        }
    };
    f32x16 acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = f32x16{};
    // Per stage: wait for its loads (the younger stages' may stay in flight)
    // and for every wave to be done with the oldest slot, start the loads of
    // the stage three ahead into it, then read this stage's fragments and run
    // its 32 MFMAs.
#pragma unroll
    for (int st = 0; st < NST - 1; ++st)
        if (st < nst) issue(st);
    // the C loads of the epilogue's first two 32-column chunks go out right
    // after the last stage's operand loads, so that the rest of the K loop
    // hides their latency (32 more loads in flight behind the stages')
    bool cpre = false;
    if (CPRE && EPI == MM_SUB && nst <= NST - 1) {
        load_chunk(0, cv[0]);
        cpre = true;
    }
    for (int st = 0; st < nst; ++st) {
        const int ahead = min(NST - 2, nst - 1 - st);   // younger stages in flight (LPS loads each)
        vm_wait(LPS * ahead + (cpre ? 16 : 0));
        __syncthreads();
        if (st + NST - 1 < nst) {
            issue(st + NST - 1);
            if (CPRE && EPI == MM_SUB && st + NST - 1 == nst - 1) {
                load_chunk(0, cv[0]);
                cpre = true;
            }
        }
        const h8* sA = lds + (st % NST) * STAGE;
        const h8* sB = sA + ABL * 64;
        // fragments double-buffered in registers: the eight LDS reads of k
        // block s + 1 are in flight while block s's 16 MFMAs run (read one
        // at a time, each read's latency stalled the single wave of a SIMD)
        h8 fa[2][4], fb[2][NJ];
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[0][i] = sA[((wm * 4 + i) * SKB) * 64 + l];
#pragma unroll
        for (int j = 0; j < NJ; ++j) fb[0][j] = sB[((wn * NJ + j) * SKB) * 64 + l];
#pragma unroll
        for (int s = 0; s < SKB; ++s) {
            const int cur = s & 1, nxt = cur ^ 1;
            if (s + 1 < SKB) {
#pragma unroll
                for (int i = 0; i < 4; ++i) fa[nxt][i] = sA[((wm * 4 + i) * SKB + s + 1) * 64 + l];
#pragma unroll
                for (int j = 0; j < NJ; ++j) fb[nxt][j] = sB[((wn * NJ + j) * SKB + s + 1) * 64 + l];
            }
            // keep the reads ahead of the MFMAs (the scheduler would otherwise
            // move each read down to its first use)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[cur][i], fb[cur][j], acc[i][j], 0, 0, 0);
        }
    }

    // Epilogue through LDS: the accumulators hold a column per lane, C is
    // column-major, so each wave writes 32 columns x 128 rows of its quarter at
    // a time into its own 17 KB of LDS (columns padded to 132 floats: the
    // 16-byte writes of 16 lanes hit distinct banks) and reads them back with
    // the lanes along the rows: every global access of a wave is two
    // contiguous 512-byte column segments.
    // accumulator layout of v_mfma_f32_32x32x16_f16: lane l holds column
    // (l & 31) and rows 8 g + 4 (l >> 5) + e of register 4 g + e
    constexpr int CP = 132;
    __syncthreads();   // every wave is done with the staging buffers
    float* ep = reinterpret_cast<float*>(lds) + w * (32 * CP);
    if (EPI == MM_SUB && !cpre) load_chunk(0, cv[0]);
    if constexpr (NCV == 2) load_chunk(1, cv[1]);
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float4 v;
                v.x = acc[i][j][4 * g];
                v.y = acc[i][j][4 * g + 1];
                v.z = acc[i][j][4 * g + 2];
                v.w = acc[i][j][4 * g + 3];
                *reinterpret_cast<float4*>(ep + (l & 31) * CP + i * 32 + g * 8 + (l >> 5) * 4) = v;
            }
        if (EPI == MM_STORE && g.tkind == 1) {
            // A tiles of C: this chunk is 4 row blocks x 2 k blocks (stores
            // have tr0 = 0)
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
                    const int kb = (col0 + j * 32) / 16 + kk;
                    if (kb >= g.tKB) continue;
                    const int rl = ib * 32 + (l & 31), kl = kk * 16 + (l >> 5) * 8;
                    h8 v;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float x = alpha * ep[(kl + e) * CP + rl];
                        const float y = tscale * x;
                        const float a = fabsf(y);
                        tmax = (a <= 3.0e38f) ? fmaxf(tmax, a) : __builtin_inff();
                        tamax = fmax(tamax, fabs((double)x));
                        if (x != x) tamax = __longlong_as_double(0x7ff8000000000000ll);
                        v[e] = (_Float16)y;
                    }
                    tout[((size_t)((row0 + ib * 32) / 32) * g.tKB + kb) * 64 + l] = v;
                }
        } else if (EPI == MM_STORE && g.tkind == 2) {
            // B tiles of C: this chunk is one column block x 8 k blocks
            const int cb = g.tcb0 + (col0 + j * 32) / 32;
#pragma unroll
            for (int kbl = 0; kbl < 8; ++kbl) {
                const int kb = row0 / 16 + kbl;
                if (kb >= g.tKB) continue;
                const int kl = kbl * 16 + (l >> 5) * 8;
                const float4 p0 = *reinterpret_cast<const float4*>(ep + (l & 31) * CP + kl);
                const float4 p1 = *reinterpret_cast<const float4*>(ep + (l & 31) * CP + kl + 4);
                const float xs[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
                h8 v;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float y = tscale * (alpha * xs[e]);
                    const float a = fabsf(y);
                    tmax = (a <= 3.0e38f) ? fmaxf(tmax, a) : __builtin_inff();
                    v[e] = (_Float16)y;
                }
                tout[((size_t)cb * g.tKB + kb) * 64 + l] = v;
            }
        }
        // subtraction with tiles: each lane puts its final values back into
        // the wave's LDS (zeros beyond M and N, as the packs pad), and the
        // tiles are cut from there after the chunk
        const bool sub_tiles = EPI == MM_SUB && g.tkind != 0 &&
                               (g.tkind == 1 ? row0 + 128 > g.tr0 : row0 < g.tKB * 16);
#pragma unroll
        for (int c2 = 0; c2 < 16; ++c2) {
            const int cl = 2 * c2 + (l >> 5);          // column within the 32
            const int col = col0 + j * 32 + cl;
            const float4 a = *reinterpret_cast<const float4*>(ep + cl * CP + rr);
            if (col >= N) {
                if (sub_tiles) *reinterpret_cast<float4*>(ep + cl * CP + rr) = float4{0.f, 0.f, 0.f, 0.f};
                continue;
            }
            float* cc = C + (size_t)col * ldc + row0 + rr;
            if (rows_ok) {
                float4 v;
                if (EPI == MM_SUB) {
                    v = cv[j % NCV][c2];
                    v.x -= alpha * a.x;
                    v.y -= alpha * a.y;
                    v.z -= alpha * a.z;
                    v.w -= alpha * a.w;
                } else {
                    v.x = alpha * a.x;
                    v.y = alpha * a.y;
                    v.z = alpha * a.z;
                    v.w = alpha * a.w;
                }
                __builtin_memcpy(cc, &v, 16);
                if (sub_tiles) *reinterpret_cast<float4*>(ep + cl * CP + rr) = v;
            } else {
                const float av[4] = {a.x, a.y, a.z, a.w};
                float fv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (row0 + rr + e < M) {
                        if (EPI == MM_SUB) fv[e] = cc[e] - alpha * av[e];
                        else fv[e] = alpha * av[e];
                        cc[e] = fv[e];
                    }
                if (sub_tiles) *reinterpret_cast<float4*>(ep + cl * CP + rr) = float4{fv[0], fv[1], fv[2], fv[3]};
            }
        }
        if (j + NCV < NJ) load_chunk(j + NCV, cv[j % NCV]);   // the next chunk into the registers chunk j freed
        if (sub_tiles && g.tkind == 1) {
            // A tiles of the result's rows from tr0 on: 4 row blocks x 2 k blocks
#pragma unroll
            for (int ib = 0; ib < 4; ++ib) {
                const int rb = row0 + ib * 32 - g.tr0;
                if (rb < 0) continue;
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
                    const int kb = (col0 + j * 32) / 16 + kk;
                    if (kb >= g.tKB) continue;
                    const int rl = ib * 32 + (l & 31), kl = kk * 16 + (l >> 5) * 8;
                    h8 v;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float y = tscale * ep[(kl + e) * CP + rl];
                        const float a = fabsf(y);
                        tmax = (a <= 3.0e38f) ? fmaxf(tmax, a) : __builtin_inff();
                        v[e] = (_Float16)y;
                    }
                    tout[((size_t)(rb / 32) * g.tKB + kb) * 64 + l] = v;
                }
            }
        } else if (sub_tiles && g.tkind == 2) {
            // B tiles of the result: one column block x 8 k blocks (rows)
            const int cb = g.tcb0 + (col0 + j * 32) / 32;
#pragma unroll
            for (int kbl = 0; kbl < 8; ++kbl) {
                const int kb = row0 / 16 + kbl;
                if (kb >= g.tKB) continue;
                const int kl = kbl * 16 + (l >> 5) * 8;
                const float4 p0 = *reinterpret_cast<const float4*>(ep + (l & 31) * CP + kl);
                const float4 p1 = *reinterpret_cast<const float4*>(ep + (l & 31) * CP + kl + 4);
                const float xs[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
                h8 v;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float y = tscale * xs[e];
                    const float a = fabsf(y);
                    tmax = (a <= 3.0e38f) ? fmaxf(tmax, a) : __builtin_inff();
                    v[e] = (_Float16)y;
                }
                tout[((size_t)cb * g.tKB + kb) * 64 + l] = v;
            }
        }
    }
    __syncthreads();   // the epilogue's LDS is the next tile's staging buffer
    if (g.tkind && g.nbat > 1) {
        // a batch's logs are per problem: merged after every tile
        pack_logs(tmax, tamax, g.maxlog ? g.maxlog + bq * g.ls : nullptr,
                  EPI == MM_STORE && g.tkind == 1 && g.amax64 ? g.amax64 + bq * g.ls : nullptr);
        tmax = 0.0f;
        tamax = 0.0;
    }
    }
    if (g.tkind && g.nbat == 1) pack_logs(tmax, tamax, g.maxlog, EPI == MM_STORE && g.tkind == 1 ? g.amax64 : nullptr);
}

// 256 x 128 tiles (two work-groups a compute unit) for every product but the
// LU's persistent bulk update; BEMHIP_MM_TN128=0 keeps 256 x 256 everywhere
static bool mm_tn128() {
    static const bool on = [] {
        const char* v = std::getenv("BEMHIP_MM_TN128");
        return !(v && v[0] == '0');
    }();
    return on;
}

// max_wg > 0: at most that many work-groups (one per compute unit), each
// looping over tiles, so that the remaining compute units stay free for
// kernels of another stream
void mm16_gemm(hipStream_t st, int epi, int tri, const void* At, int Ka, const void* Bt, int Kb, int c0, int M, int N,
               int K, float* C, long ldc, float alpha, int max_wg, const Mm16Tiles* tiles, const Mm16Bat* bat,
               const Mm16Seg* seg) {
    if (M <= 0 || N <= 0) return;
    BH_CHECK(c0 % MM_TILE == 0, BEMHIP_ERR_ARG, "mm16: column offset must be a multiple of 256");
    BH_CHECK(K <= Ka && K <= Kb, BEMHIP_ERR_ARG, "mm16: K exceeds the operand tiles");
    BH_CHECK(!(tiles && epi == MM_SUB && tiles->amax64), BEMHIP_ERR_ARG, "mm16: no max |x| for subtractions");
    BH_CHECK(!(tiles && epi != MM_SUB && tiles->r0 != 0), BEMHIP_ERR_ARG, "mm16: row offset only for subtractions");
    BH_CHECK(!(tiles && tiles->r0 % MM_TILE != 0), BEMHIP_ERR_ARG, "mm16: tile row offset must be a multiple of 256");
    MmArgs g{};
    g.At = static_cast<const h8*>(At);
    g.Bt = static_cast<const h8*>(Bt);
    g.C = C;
    g.ldc = ldc;
    g.alpha = alpha;
    g.KBa = mm_pad(Ka, MM_BK) / 16;
    g.KBb = mm_pad(Kb, MM_BK) / 16;
    g.cb0 = c0 / 32;
    g.M = M;
    g.N = N;
    g.nks = mm_pad(K, MM_BK) / MM_BK;
    if (seg) {
        BH_CHECK(epi == MM_SUB && tri == 0 && !bat, BEMHIP_ERR_ARG, "mm16: a second segment only for single updates");
        BH_CHECK(K % MM_BK == 0 && seg->K > 0 && seg->K <= seg->Ka && seg->K <= seg->Kb && seg->c0 % MM_TILE == 0,
                 BEMHIP_ERR_ARG, "mm16: bad second segment");
        g.At2 = static_cast<const h8*>(seg->At);
        g.Bt2 = static_cast<const h8*>(seg->Bt);
        g.KBa2 = mm_pad(seg->Ka, MM_BK) / 16;
        g.KBb2 = mm_pad(seg->Kb, MM_BK) / 16;
        g.cb02 = seg->c0 / 32;
        g.nks1 = g.nks;
        g.nks += mm_pad(seg->K, MM_BK) / MM_BK;
    }
    const bool bulk = seg && max_wg > 0;
    const int tn = bulk || !mm_tn128() ? 256 : 128;
    g.tiles_m = (M + MM_TILE - 1) / MM_TILE;
    g.tiles_n = (N + tn - 1) / tn;
    g.nbat = 1;
    if (bat) {
        BH_CHECK(bat->n >= 1 && bat->sa % 16 == 0 && bat->sb % 16 == 0 && bat->st % 16 == 0, BEMHIP_ERR_ARG,
                 "mm16: bad batch");
        g.nbat = bat->n;
        g.sAt = bat->sa / 16;
        g.sBt = bat->sb / 16;
        g.sC = bat->sx;
        g.sT = bat->st / 16;
        g.ebat = bat->e;
        g.ek = bat->ek;
        g.ek2 = bat->ek2;
        g.ls = bat->ls;
    }
    g.per_xcd = (int)(((long long)g.tiles_m * g.tiles_n * g.nbat + 7) / 8);
    // max_wg counts compute units: two 256 x 128 work-groups share one
    const int wg_cap = tn == 128 ? 2 * max_wg : max_wg;
    g.wg_xcd = max_wg > 0 ? std::max(1, std::min(g.per_xcd, wg_cap / 8)) : g.per_xcd;
    if (tiles) {
        g.tout = static_cast<h8*>(tiles->out);
        g.tkind = tiles->b_operand ? 2 : 1;
        g.tKB = mm_pad(tiles->k, MM_BK) / 16;
        g.tcb0 = tiles->c0 / 32;
        g.tscale = tiles->scale;
        g.maxlog = tiles->maxlog;
        g.amax64 = tiles->amax64;
        g.tr0 = tiles->b_operand ? 0 : tiles->r0;
        BH_CHECK(tiles->c0 % 32 == 0, BEMHIP_ERR_ARG, "mm16: tile column offset must be a multiple of 32");
    }
    const dim3 grid(8 * g.wg_xcd);
    if (epi == MM_SUB) {
        BH_CHECK(tri == 0, BEMHIP_ERR_ARG, "mm16: triangular operands only for stores");
        if (bulk) k_mm16<MM_SUB, 0, true><<<grid, 256, 0, st>>>(g);
        else if (tn == 128) k_mm16<MM_SUB, 0, false, 128><<<grid, 256, 0, st>>>(g);
        else k_mm16<MM_SUB, 0><<<grid, 256, 0, st>>>(g);
    } else if (tri == 1) {
        if (tn == 128) k_mm16<MM_STORE, 1, false, 128><<<grid, 256, 0, st>>>(g);
        else k_mm16<MM_STORE, 1><<<grid, 256, 0, st>>>(g);
    } else if (tri == 2) {
        if (tn == 128) k_mm16<MM_STORE, 2, false, 128><<<grid, 256, 0, st>>>(g);
        else k_mm16<MM_STORE, 2><<<grid, 256, 0, st>>>(g);
    } else {
        if (tn == 128) k_mm16<MM_STORE, 0, false, 128><<<grid, 256, 0, st>>>(g);
        else k_mm16<MM_STORE, 0><<<grid, 256, 0, st>>>(g);
    }
    BH_HIP(hipGetLastError());
}

}  // namespace bemhip
