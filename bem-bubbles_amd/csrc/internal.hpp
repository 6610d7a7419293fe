// Internal declarations of libbemhip (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bemhip.h"

namespace bemhip {

// ---------------------------------------------------------------- errors --
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BH_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::bemhip::Error(e_ == hipErrorOutOfMemory ? BEMHIP_ERR_OOM : BEMHIP_ERR_HIP, \
                                  std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

#define BH_BLAS(expr)                                                                 \
    do {                                                                              \
        rocblas_status s_ = (expr);                                                   \
        if (s_ != rocblas_status_success)                                             \
            throw ::bemhip::Error(BEMHIP_ERR_HIP, std::string(#expr) + ": " +         \
                                                      rocblas_status_to_string(s_));  \
    } while (0)

#define BH_CHECK(cond, code, msg)                          \
    do {                                                   \
        if (!(cond)) throw ::bemhip::Error((code), (msg)); \
    } while (0)

// ------------------------------------------------------- device buffers --
// device allocations made by DevBuf so far (a captured graph is only replayed
// while no buffer it may address has been reallocated since its capture)
extern std::atomic<unsigned long long> g_devbuf_allocs;

// Grow-only device allocation (contents are not preserved on growth).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void* ensure(size_t n) {
        if (n <= bytes && p) return p;
        release();
        if (n == 0) n = 256;
        BH_HIP(hipMalloc(&p, n));
        bytes = n;
        g_devbuf_allocs.fetch_add(1);
        return p;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ----------------------------------------------------- mesh topology --
// Host-built, topology-only tables (rebuilt when the connectivity changes).
constexpr int PATCH = 12;          // columns per assembly patch (12: 6 waves per SIMD; measured best of 8-16)
constexpr int ASM_WG_TARGET = 32768;   // assembly work-groups per launch (measured: 8192 28.2 ms, 32768 25.9 ms at N = 40962)
constexpr int ASM_ROWS = 128;      // collocation rows per assembly work-group

// One dense BEM problem inside a mesh: all vertices of the mesh (coupled
// bubbles, the reference's semantics) or one loose part (independent bubbles).
// Its G block is n x n, column-major with leading dimension n, at element
// offset gofs of the G buffer; its vertices are [v0, v0 + n).
struct AsmProb {
    int32_t v0, n;
    int64_t gofs;
    int32_t groups;   // patch groups = H phi partial slices of its rows
    int32_t p0, p1;   // its patches
    int32_t pad;
};
// elements between the G blocks of two problems of n unknowns laid out one
// after the other (256-B aligned in FP32 and FP64)
inline int64_t gblock_stride(int n) { return ((int64_t)n * n + 63) / 64 * 64; }
// One assembly work-group: 128 rows of a problem x a range of its patches
struct AsmItem {
    int32_t prob, row0, p0, p1, g, pad0, pad1, pad2;
};

struct Topology {
    int64_t N = 0, M = 0;
    int npatch = 0;
    bool independent = false;
    std::vector<AsmProb> probs;
    std::vector<AsmItem> items;
    int max_groups = 0;
    std::vector<uint32_t> tri;             // M x 3
    // assembly patches: vertex -> (patch, slot); per patch, the incident triangles
    std::vector<int32_t> patch_col;        // npatch * PATCH vertex ids (-1 = empty slot)
    std::vector<int32_t> patch_ptr;        // npatch + 1
    std::vector<uint32_t> patch_tri;       // entries: triangle id
    std::vector<uint32_t> patch_slots;     // entries: 3 x 8-bit slot (0xFF = other patch)
    // vertex -> incident triangles (ascending), Mesh.cpp:75-84
    std::vector<int32_t> vptr, vtri;
    // loose parts (Mesh.cpp:382-440): per vertex part id; per part triangle list
    int nparts = 0;
    std::vector<int32_t> vpart;
    std::vector<int32_t> part_tptr, part_tri;   // triangles of each loose part (ascending)
    std::vector<int32_t> part_vptr, part_vtx;   // vertices of each loose part (ascending)
    std::vector<int32_t> part_pptr;        // per part, its patch range
    double redundancy = 0.0;               // patch entries / M
};

// independent = every loose part is its own problem (parts must then be
// contiguous vertex ranges in increasing order)
Topology build_topology(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, bool independent = false);

// Device copies of the topology tables
struct DevTopology {
    uint64_t gen = 0;                  // distinct for every upload (in the process)
    int64_t N = 0, M = 0;
    int npatch = 0, nparts = 0, nentries = 0, nitems = 0, max_groups = 0;
    bool independent = false;
    std::vector<AsmProb> probs;        // host copy (solve loop)
    int64_t gsize = 0;                 // elements of all G blocks
    DevBuf tri, patch_col, patch_ptr, patch_tri, patch_slots, vptr, vtri, vpart, part_tptr, part_tri, part_vptr,
        part_vtx;
    DevBuf d_probs, d_items;
    void upload(const Topology& t, hipStream_t s);
};

// ------------------------------------------------------------- context --
// small problems of a batch (see small_solve_launch)
constexpr int SMALL_MAX_N = 1024;
struct SmallJob {
    int64_t gofs;     // element offset of the problem's G block
    int32_t n, v0;    // unknowns, first vertex (offset into b and x)
};

struct ClassWork;

// what a captured batch stage depends on besides the buffers' contents
struct BatchGraphKey {
    uint64_t topo_gen;
    const double* xyz;
    const double* phi;
    double* hphi;
    double* psi;
    double* G64;
    float* G32;
    uint32_t flags;
    double tol;
    int maxit, nb16, split, lookahead, sweeps, reserve, small_n;
    bool operator==(const BatchGraphKey& o) const {
        return topo_gen == o.topo_gen && xyz == o.xyz && phi == o.phi && hphi == o.hphi && psi == o.psi &&
               G64 == o.G64 && G32 == o.G32 && flags == o.flags && tol == o.tol && maxit == o.maxit &&
               nb16 == o.nb16 && split == o.split && lookahead == o.lookahead && sweeps == o.sweeps &&
               reserve == o.reserve && small_n == o.small_n;
    }
};
struct BatchGraph {
    BatchGraphKey key;
    hipGraphExec_t exec = nullptr;
    unsigned long long allocs = 0;   // g_devbuf_allocs at capture
    int captures = 0, launches = 0;
};

struct LuWork {
    DevBuf ipiv, panel_ws, flag, inv_lo, inv_up, backup, maxl, diag;   // maxl: max |l|, then top-block sync words
    // per outer step: inverse of the panel's unit-lower diagonal block and its
    // L21 (two of each, alternating by step parity), inverse of U11, U12
    DevBuf winv_l[4], l21[4], winv_u, u12;   // four parities: the fp16 LU's merged steps keep two panel pairs
    int inv_n = 0;   // N of the factor whose diagonal-block inverses are current
    const void* inv_src = nullptr;   // and that factor
    DevBuf verdict;   // merged fp16 LU: max |l| over its panels and their error words
    // fp16 LU (mm16): fragment tiles of A21, U11^-1, L11^-1 and L21 (the last
    // two alternating by step parity), of the block row A12 and of U12; the
    // max |scaled value| of every pack (float bits)
    DevBuf t_a21, t_wu, t_wl[4], t_l21[4], t_a12, t_u12[4], f16log;
    DevBuf tinv;   // work of the triangular inverses (2 tri_inv_tmp_elems(NB) elements)
};

// Everything one dense solve needs: its streams, BLAS handles, work buffers,
// pinned host words and instrumentation.  The context is one solver (lane 0);
// a batch of independent problems adds more, each driven by its own host
// thread so the problems' latency-bound kernels overlap on the device.
struct Solver {
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    hipStream_t stream2 = nullptr;      // LU look-ahead: panels of the next block column
    hipStream_t stream3 = nullptr;      // fp16 LU with merged updates: the narrow products
    hipEvent_t ev_m[4] = {};            // its hand-overs: narrow products done, big update done
    rocblas_handle blas2 = nullptr;
    // look-ahead only while the trailing update of a step has at least this
    // many columns besides the next panel's (0: always; measured best at
    // N = 40962 in FP64 and mixed precision)
    int la_min_rest = 0;
    hipEvent_t ev_la[4] = {};   // LU: next panel's columns ready, panel done, max |l| copied, L21 copied
    bool lookahead = true;
    bool speculate = true;              // panels without interchanges, verified against GEPP's choice
    bool np_top = true;                 // a panel's top block in one persistent launch (else 32-column steps)
    bool np_top32 = true;               // FP32: that launch with the stripes held in registers
    unsigned int np_epoch = 0;          // publication stamp of the persistent top-block launches
    int nb64 = 768, nb32 = 512, nb_split = 1024;   // LU outer block per precision (measured at N = 40962)
    int nb16 = 1024;                               // fp16 LU (a multiple of 256)
    // fp16 LU: compute units the trailing update leaves to the panel stream
    // (its launch is persistent on the others)
    int mm_reserve = 32;
    int ncu = 256;   // compute units of the device
    int panels = 0, panels_pivoted = 0;
    bool trsv_persistent = true;
    // test hook (BEMHIP_TRSV_FORCE_GIVEUP=1): the host pre-sets the persistent
    // triangular solve's error word, so every waiting work-group gives up at
    // once and the solve must report BEMHIP_ERR_HIP
    bool trsv_force_giveup = false;
    // test hook (BEMHIP_BATCH_FORCE_REDO=q): problem q of every size class of a
    // batched fp16 LU is solved again on its own as if its check had failed
    int batch_force_redo = -1;
    // fp16 LU: the next panel's A21 tiles were written by the update that
    // produced A21 (set by the main thread before it hands the panel over)
    bool lu16_a21_ready = false;
    // single problems of at most SMALL_MAX_N unknowns through the one-work-group
    // FP64 solver (BEMHIP_SMALL_SINGLE=1; leaves no factors for a transposed solve)
    bool small_single = false;
    DevBuf small1;   // its job record and info word
    unsigned int trsv_epoch = 0;
    // split-fp16 updates with one scale for every U12 block of a factorisation
    // (a host-side GEMM alpha); each block's max is logged and checked after
    // the factorisation, which is redone with per-block scales if one fell
    // outside the window the fixed scale covers
    bool split_fixed = true;
    // fp16 LU: every product of the factorisation on fp16 matrix cores with
    // FP32 accumulation (mm16); operands scaled by powers of two from max |diag|
    // (2^lu16_e), an overflow of the fp16 range detected after the
    // factorisation (lu16_overflow), which the solver then redoes in FP32
    bool lu16 = false, lu16_overflow = false;
    int lu16_e = 0, lu16_redos = 0;
    // fp16 LU: the trailing update of two panels in one launch (BEMHIP_LU16_MERGE=0:
    // step by step); lu16_unmerged counts factorisations redone step by step
    // because a panel needed row interchanges
    bool lu16_merge = true;
    // merged fp16 LU: pairs with at least this many trailing rows finish
    // U12 before the next panel's columns (the bulk update goes first;
    // measured at N = 40962: 25000 131.1-131.5 ms per stage, 0 131.7-132.3,
    // never 132.0)
    long long lu16_bulk_first = 25000;
    int lu16_unmerged = 0;
    bool split_fixed_on = false, split_redo = false;
    int split_eb = 0, split_calls = 0, split_nlog = 0;
    int split_redos = 0;   // factorisations redone with per-block scales (lifetime count)
    DevBuf split_log;
    // workspace
    DevBuf vec[12], scal, gemv_part;
    LuWork lu;
    DevBuf split_a[2], split_b, split_mx, perm;   // split_a alternates by LU step parity
    std::vector<int> h_ipiv, h_perm;
    double* host_scal = nullptr;  // pinned
    // instrumentation of trailing updates
    bool timing = true;
    std::vector<hipEvent_t> gemm_ev;   // start/stop pairs of the trailing updates of one factorisation
    std::vector<double> gemm_fl;
    size_t gemm_ev_used = 0;
    double gemm_ms = 0.0, gemm_flops = 0.0;
    int gemm_launches = 0;

    // streams, handles, pinned words and events on the current device; the
    // tuning knobs are copied from `cfg` when given
    void init(const Solver* cfg);
    void release();
};

}  // namespace bemhip

struct bemhip_ctx : bemhip::Solver {
    int device = 0;
    std::string err;
    int cus = 0;
    // assembly and per-problem buffers
    bemhip::DevBuf triq, part_hphi, part_hrow, hphi, xyz, phi, G64, G32, H;
    bemhip::DevBuf ext_rec, ext_part, ext_tgt, ext_out, ext_tri, ext_psi;   // exterior potential
    double* ext_scratch = nullptr;   // this device's address of the no-allocation scratch array
    bemhip::DevBuf adj_coef, adj_tc, adj_tpart, adj_rpart, adj_sc, adj_zero;   // adjoint partials
    hipEvent_t ev[4] = {};
    // extra solvers for batches of independent problems (created on first use)
    std::vector<std::unique_ptr<bemhip::Solver>> lanes;
    int max_lanes = 8;
    // batches: problems of at most small_n unknowns (and LU solvers) go to the
    // one-launch small-problem solver, enqueued last on the least loaded
    // solver lane (0: never)
    int small_n = bemhip::SMALL_MAX_N;
    hipEvent_t small_ev[2] = {};   // start, end of the launch
    std::vector<bemhip::SmallJob> small_jobs_h;
    bemhip::DevBuf small_jobs, small_info;
    // batches: problems of more than small_n unknowns solved by the lu16_ir64
    // solver are factored in lock step per size class (BEMHIP_BATCH_LU=0: on
    // the solver lanes instead); refinement sweeps enqueued before the first
    // host read-back of a class
    bool batch_lu = true;
    int batch_sweeps = 2;
    // each size class split into this many sub-classes on lanes of their own
    // (their panels and updates then overlap without a fixed split of the
    // compute units); look-ahead inside a (sub-)class
    int batch_split = 1;
    bool batch_lookahead = true;
    std::vector<std::unique_ptr<bemhip::ClassWork>> class_work;
    std::vector<bemhip::SmallJob> small_jobs_up;   // the job table last uploaded
    const bemhip::SmallJob* small_jobs_dev = nullptr;
    // lu16_ir64 batches: a stage's device work captured into a graph the
    // second time its key comes by, and replayed.  Off: with it on, RK4
    // stepping of a batch ended in a host segmentation fault whose cause has
    // not been found (BEMHIP_BATCH_GRAPH=1 switches it on)
    bool batch_graph = false;
    std::vector<bemhip::BatchGraph> batch_graphs;
};

namespace bemhip {

// ------------------------------------------------------------ assembly --
struct AsmOut {
    double* G64;     // may be null only when G32 is given
    float* G32;      // optional FP32 copy
    double* H;       // optional
    double* hphi;    // optional (needs phi)
};
void assemble_device(bemhip_ctx* ctx, const DevTopology& topo, const double* d_xyz, const double* d_phi,
                     uint32_t flags, const AsmOut& out);

// ------------------------------------------- fp16 matrix-core products --
// (mm16.hip) fp16 fragment-tiled operands: blocks of 32 rows (or columns) x 16 k
// in MFMA operand order, padded to MM_TILE rows and MM_BK k
constexpr int MM_TILE = 256;
constexpr int MM_BK = 64;
constexpr int MM_SUB = 0, MM_STORE = 1;
inline int mm_pad(int n, int q) { return (n + q - 1) / q * q; }
// elements allocated past the end of an FP32 matrix that a trailing update
// (MM_SUB) works on: with fewer than 4 rows its clamped C reads may touch up
// to 3 elements past the block's last row
constexpr size_t MM_C_TAIL = 64;
// bytes of the tiles of R rows (or columns) x K
inline size_t mm_tile_bytes(int R, int K) { return (size_t)mm_pad(R, MM_TILE) * mm_pad(K, MM_BK) * 2; }
// fp16 tiles of scale * X; kcontig: X is K x R column-major (B operand), else R x K
// (A operand); copy (optional, B only, ld ldc) receives X at the same positions;
// maxlog (optional) the max |scale x| as float bits, amax64 (optional, A only)
// max |x| as the bits of a non-negative double
// A batch of equally shaped products (or packs): problem b's operands start at
// the given pointer plus b times the stride of that operand, and its scales
// carry the factor 2^(ek e[b]) (e: per-problem exponents on the device; null
// for none).  The logs (maxlog, amax64) are per problem, ls elements apart.
struct Mm16Bat {
    int n = 1;                    // problems
    long long sx = 0;             // pack: source X / gemm: C, in floats
    long long sa = 0, sb = 0;     // pack: output tiles (sa) / gemm: A and B tiles, in bytes
    long long st = 0;             // gemm: output tiles (Mm16Tiles::out), in bytes
    const int* e = nullptr;
    int ek = 0;                   // pack: scale / gemm: alpha
    int ek2 = 0;                  // gemm: the output tiles' scale
    int ls = 0;
};
void mm16_pack(hipStream_t st, bool kcontig, int R, int K, const float* X, long ld, float scale, void* tiles,
               unsigned int* maxlog, float* copy = nullptr, long ldc = 0, unsigned long long* amax64 = nullptr,
               const Mm16Bat* bat = nullptr);
// C (M x N, ld ldc) -= alpha A B (MM_SUB) or = alpha A B (MM_STORE) over the first
// K of the k extent; A: tiles of M x Ka, B: tiles of Kb x (c0 + N) read from
// column c0 (a multiple of MM_TILE); tri 1: A lower, 2: B upper triangular;
// max_wg > 0: a persistent launch of at most max_wg work-groups (one per CU);
// tiles (MM_STORE only): the stored C also as fp16 tiles (as mm16_pack would
// make them from C)
struct Mm16Tiles {
    void* out;                   // tile buffer
    bool b_operand;              // B tiles (k = row of C) or A tiles (k = column of C)
    int k;                       // k extent of the buffer (rows of C for B, columns for A)
    int c0;                      // B: column of C's first column within the buffer (multiple of 32)
    float scale;
    unsigned int* maxlog;        // max |scale x| (float bits)
    unsigned long long* amax64;  // A only: max |x| (double bits)
    int r0 = 0;                  // A only: C's rows from r0 on (a multiple of 256) are the buffer's rows
};
// a second operand segment of an update: C -= alpha (A B + A2 B2), A2 the
// tiles of M rows x Ka (first K of them used), B2 of Kb x (c0 + N) from column
// c0 (a multiple of MM_TILE); the first segment's K must be a multiple of MM_BK
struct Mm16Seg {
    const void* At;
    int Ka;
    const void* Bt;
    int Kb, c0, K;
};
void mm16_gemm(hipStream_t st, int epi, int tri, const void* At, int Ka, const void* Bt, int Kb, int c0, int M, int N,
               int K, float* C, long ldc, float alpha, int max_wg = 0, const Mm16Tiles* tiles = nullptr,
               const Mm16Bat* bat = nullptr, const Mm16Seg* seg = nullptr);

// ---------------------------------------------------------- dense LA --
// LU with partial pivoting (column-major, lda = N), pivots as absolute rows.
// potential at P points off the surface (compute_exterior_pot, LinLinSim.cpp:440-477);
// all pointers on the device, out[P]
void exterior_device(bemhip_ctx* ctx, int64_t M, const double* xyz, const uint32_t* tri, const double* phi,
                     const double* psi, int64_t P, const double* tgt, double* out);
// the same with every work buffer in a fixed device array (no allocation);
// validates the triangle indices and finiteness on the device; synchronous
void exterior_device_noalloc(bemhip_ctx* ctx, int64_t N, int64_t M, const double* xyz, const uint32_t* tri,
                             const double* phi, const double* psi, int64_t P, const double* tgt, double* out);
// this device's address of that array into ctx (once, at context creation)
void exterior_scratch_init(bemhip_ctx* ctx);
// Wl, Wu (n x n, ld n) = inverses of the unit-lower and the non-unit upper
// triangle of A (ld lda), each with its other triangle written as zeros, in
// the same launches (a null W skips its triangle); tmp_l and tmp_u hold
// tri_inv_tmp_elems(n) elements each (trinv.hip)
// A batch of problems for tri_inv2: problem q's A, W and tmp start q sA, q sW,
// q sT elements further
struct TriBat {
    int n;
    int64_t sA, sW, sT;
};
template <class T>
void tri_inv2(hipStream_t st, int n, const T* A, int lda, T* Wl, T* Wu, T* tmp_l, T* tmp_u,
              const TriBat* bat = nullptr);
size_t tri_inv_tmp_elems(int n);
template <class T> int lu_factor(Solver* ctx, int N, T* A, int* ipiv, bool split16 = false);
// x (double, device) <- (LU)^-1 P x with LU in precision T
template <class T> void lu_solve(Solver* ctx, int N, const T* LU, const int* ipiv, double* x);
// after a factorisation: diagonal-block inverses for the persistent solves
template <class T> void lu_prepare_solve(Solver* ctx, int N, const T* LU);
// r = b - A x (A FP64 column-major), or y = A x when b == nullptr
void gemv64(Solver* ctx, int N, const double* A, const double* x, const double* b, double* r);
// solver front end on a device-resident FP64 G (destroyed by LU64) and rhs b.
void solve_device(Solver* ctx, int N, double* G64, float* G32, const double* b, double* x,
                  int solver, double tol, int maxit, bemhip_solve_stats* st, bool report_residual = true);
// G^T x = b with the factors left by the last LU-based solve_device call on
// this context (same N and solver; G64 is the unfactored FP64 G for the mixed
// solvers, whose transposed solve is refined against it)
void solve_transposed_device(Solver* ctx, int N, const double* G64, const float* G32, int solver,
                             const double* b, double* x, double tol, int maxit, bemhip_solve_stats* st);
// ---------------------------------------- batched fp16 LU (size classes) --
// A size class of a batch: cnt independent problems of n unknowns whose G
// blocks (FP64 and the FP32 copy) lie gblock_stride(n) elements apart from
// element offset gofs0; problem q's right-hand side and solution start at
// v0[q] of b and x.
struct ClassJob {
    int n = 0, cnt = 0;
    int64_t gofs0 = 0;
    std::vector<int32_t> v0;
    // the next panel of every problem factored on the solver's second stream
    // while the update of the current step runs (else one stream, every
    // launch free to take the whole device)
    bool lookahead = true;
};
// Device workspace of one class (grown on demand, kept between stages).
struct ClassWork {
    int n = 0, cnt = 0, nb = 0, nbk = 0, nch = 0;
    size_t tb_a21 = 0, tb_w = 0, tb_a12 = 0;   // bytes of a problem's A21 / NB x NB / block-row tiles
    DevBuf e, top_sync, tickets, maxl64, f16log, linv, winv_l[2], winv_u, tinv, t_a21, t_l21[2], t_wu, t_wl[2], t_a12,
        t_u12;
    DevBuf inv_lo, inv_up, tsync, vecs, part, scal, v0;
    unsigned int top_epoch = 0, trsv_epoch = 0;
    std::vector<int32_t> v0_h;   // host copy of the uploaded offsets
    bool timed = false;          // the last enqueue recorded its timing events
    hipEvent_t ev[5] = {};   // next panel's columns ready, panel done, factorisation start, end, sweeps end
    float ms_factor = 0.0f;
    ClassWork() = default;
    ClassWork(const ClassWork&) = delete;
    ClassWork& operator=(const ClassWork&) = delete;
    ~ClassWork();
};
// Enqueues, on the solver's two streams and without any host
// synchronisation, the lock-step fp16 LU of every problem of the class (one
// launch per block product over all of them), the FP32 triangular solves and
// `sweeps` FP64 refinement sweeps, each of which skips the problems whose
// refinement has already stopped, and writes the solutions to x.  Nothing in
// it reads the host or allocates once the class is prepared, so it may be
// captured into a graph (events = false: no timing events).  Panels are factored without row
// interchanges; a problem whose multipliers exceed 1 in magnitude (partial
// pivoting would have interchanged rows) or whose fp16 operands overflowed is
// flagged for class_lu16_finish.
void class_lu16_enqueue(Solver* ctx, ClassWork* w, const ClassJob& cj, double* G64, float* G32, const double* b,
                        double* x, double tol, int sweeps, bool events = true);
// sizes the class's workspace and uploads its offsets (host synchronous; called
// by class_lu16_enqueue, and before a stream capture so that none of this
// happens inside it)
void class_lu16_prepare(Solver* ctx, ClassWork* w, const ClassJob& cj);
// Reads the class's status back, runs further sweeps for problems still
// refining (at most maxit in all), writes the solutions to x and solves every
// flagged problem again on its own with the single-problem solver (partial
// pivoting or FP32 products from a fresh FP32 copy of G).
void class_lu16_finish(Solver* ctx, ClassWork* w, const ClassJob& cj, double* G64, float* G32, const double* b,
                       double* x, double tol, int maxit, bemhip_solve_stats* st);

// small problems of a batch: one work-group each, FP64 GEPP + solve in place
// on the problem's G block (destroyed), all in one launch; info[q] = LAPACK info
void small_solve_launch(hipStream_t st, const SmallJob* d_jobs, int count, double* G64, const double* b, double* x,
                        int* d_info);
// device reductions
double dev_absmax(Solver* ctx, int n, const double* x);
double dev_dot(Solver* ctx, int n, const double* x, const double* y);

// ------------------------------------------------------------- stage --
// scalars of the Bernoulli right-hand side; V_0 and the forcing are per part
// (an array of [V_0, amplitude, frequency, phase] per loose part)
struct StageParams {
    double p_inf, epsilon, sigma, gamma;
};
void tri_geometry(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* pot,
                  double* tnorm, double* tgrad, double* tvol);
void vertex_normals(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, double* vn);
void tri_curvature(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* vn,
                   double* tkappa);
void part_volumes(bemhip_ctx* ctx, const DevTopology& topo, const double* tvol, double* pvol);
void vertex_velocity(bemhip_ctx* ctx, const DevTopology& topo, const double* tnorm, const double* tgrad,
                     const double* psi, double* vel);
void vertex_dpot(bemhip_ctx* ctx, const DevTopology& topo, const StageParams& sp, const double* xyz,
                 const double* vel, const double* tkappa, const double* pvol, const double* pphys, double t,
                 double* dpot);
void axpy_state(bemhip_ctx* ctx, int64_t N, double c, const double* kx, const double* x1, double* xs,
                const double* kp, const double* p1, double* ps);
// x = c dx (+ x), p = c dp (+ p) over N vertices (x: N x 3)
void axpy_pair(bemhip_ctx* ctx, int64_t N, double c, const double* dx, double* x, const double* dp, double* p,
               bool accumulate);
void rk4_final(bemhip_ctx* ctx, int64_t N, double dt, const double* kx, const double* kp,
               const double* x1, double* x, double* phi, double* avg, int wall_guards, double wall_eps);
void psi_project(bemhip_ctx* ctx, int64_t N, const double* avg, const double* vn, double* psi);
void euler_update(bemhip_ctx* ctx, int64_t N, double dt, const double* vel, const double* dpot,
                  double* x, double* phi);
void dt_reduce(bemhip_ctx* ctx, int64_t N, const double* vel, const double* dpot, double* out2);
int any_negative_x(bemhip_ctx* ctx, int64_t N, const double* xyz);
// pinned contact line (ColocSimPin): the last npin vertices
void pin_copy_tail(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* src, double* dst);
void pin_zero_tail(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* pot, double* pz);
void pin_swap_columns(bemhip_ctx* ctx, int64_t N, int64_t npin, double* G, float* G32, const double* H,
                      const double* pot, double* rhs);
void pin_drift(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* vn, const double* psil, double* vel,
               double* pot);

// ----------------------------------------------------------- adjoint --
// forward quantities of one stage that its reverse sweep reads
struct AdjStageIn {
    const double* xyz;     // stage positions
    const double* pot;     // stage potential
    const double* psi;     // stage dphi/dn
    const double* vel;     // stage vertex velocities
    const double* pvol;    // part volumes at xyz
    const double* pphys;   // per-part [V_0, amplitude, frequency, phase]
    double epsilon, gamma;
};
// kxbar += d(kp)/d(kx)^T kpbar; psibar = cotangent of psi; per-triangle
// partials of the velocity / volume terms kept in the context
void adj_stage_pre(bemhip_ctx* ctx, const DevTopology& topo, const AdjStageIn& in, double* kxbar, const double* kpbar,
                   double* psibar);
// with lam = G^-T psibar: xbar, pbar += the stage's input cotangents
void adj_stage_post(bemhip_ctx* ctx, const DevTopology& topo, const AdjStageIn& in, const double* lam, double* xbar,
                    double* pbar);
// xbar += vcoef[part] * d(part volume)/dx
void adj_volume(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* vcoef, double* xbar);

}  // namespace bemhip
