// C-ABI of libbemhip and the device-resident simulation handle that mirrors
// the reference's LinLinSim / ColocSim state machine (LinLinSim.cpp:26-284,
// Simulation.hpp:29-167).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <thread>
#include <exception>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "internal.hpp"
#include "quad_rules.hpp"

using namespace bemhip;

namespace bemhip {

std::atomic<unsigned long long> g_devbuf_allocs{0};

__global__ void k_d2f(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
}  // namespace bemhip

struct bemhip_sim {
    bemhip_ctx* ctx = nullptr;
    bemhip_sim_params P{};
    Topology th;
    DevTopology topo;
    int64_t N = 0, M = 0;
    double time = 0.0;
    DevBuf x, phi, psi, x1, p1, xs, ps, kx, kp, avg, vn, tnorm, tgrad, tvol, pvol, tk, hphi, psil, pphys;
    int64_t n_pin = 0;   // ColocSimPin: the last n_pin vertices are pinned on the wall
    DevBuf pz;           // stage potential with the pinned entries zeroed
    // per-part V_0, field amplitude, frequency, phase; NaN = the scalar in P
    std::vector<double> part_v0, part_amp, part_freq, part_phase;
    // adjoint work space: start state, checkpoints, segment states, cotangents
    DevBuf a_x0, a_p0, a_ckx, a_ckp, a_sgx, a_sgp, a_kxb, a_kpb, a_xb, a_pb, a_xsb, a_psb, a_psib, a_lam, a_vc, a_kxt,
        a_kpt;
    bemhip_solve_stats last{};
    bemhip_solve_stats accum{};
};

namespace {

template <class F> int guarded(bemhip_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->err.clear();
        return BEMHIP_OK;
    } catch (const Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "host allocation failed";
        return BEMHIP_ERR_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return BEMHIP_ERR_HIP;
    }
}

void check_finite(const double* p, int64_t n, const char* what) {
    if (!p) return;
    for (int64_t i = 0; i < n; ++i)
        BH_CHECK(std::isfinite(p[i]), BEMHIP_ERR_NAN, std::string("non-finite value in ") + what);
}

template <class T> T* dev(DevBuf& b, size_t n) { return static_cast<T*>(b.ensure(sizeof(T) * std::max<size_t>(n, 1))); }

void h2d(void* d, const void* h, size_t bytes, hipStream_t s) {
    BH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
}
void d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
    BH_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    BH_HIP(hipStreamSynchronize(s));
}

bool mixed(int solver) {
    return solver == BEMHIP_SOLVE_LU32_IR64 || solver == BEMHIP_SOLVE_LU16X3_IR64 || solver == BEMHIP_SOLVE_LU16_IR64;
}

float event_ms(hipEvent_t a, hipEvent_t b) {
    BH_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void merge_stats(bemhip_solve_stats& sum, const bemhip_solve_stats& one) {
    sum.iterations = std::max(sum.iterations, one.iterations);
    sum.converged = sum.converged && one.converged;
    sum.residual = std::max(sum.residual, one.residual);
    sum.correction = std::max(sum.correction, one.correction);
    sum.ms_factor += one.ms_factor;
    sum.ms_solve += one.ms_solve;
    sum.gemm_ms += one.gemm_ms;
    sum.gemm_flops += one.gemm_flops;
    sum.gemm_launches += one.gemm_launches;
    sum.panels += one.panels;
    sum.panels_pivoted += one.panels_pivoted;
}

// the small problems' job table on the device (uploaded when it changed)
static SmallJob* small_upload(bemhip_ctx* ctx, hipStream_t ls) {
    const int nsmall = (int)ctx->small_jobs_h.size();
    SmallJob* dj = static_cast<SmallJob*>(ctx->small_jobs.ensure(sizeof(SmallJob) * nsmall));
    ctx->small_info.ensure(sizeof(int) * nsmall);
    const bool same = ctx->small_jobs_up.size() == ctx->small_jobs_h.size() &&
                      std::memcmp(ctx->small_jobs_up.data(), ctx->small_jobs_h.data(), sizeof(SmallJob) * nsmall) == 0;
    if (!same || ctx->small_jobs_dev != dj) {
        BH_HIP(hipMemcpyAsync(dj, ctx->small_jobs_h.data(), sizeof(SmallJob) * nsmall, hipMemcpyHostToDevice, ls));
        BH_HIP(hipStreamSynchronize(ls));
        ctx->small_jobs_up = ctx->small_jobs_h;
        ctx->small_jobs_dev = dj;
    }
    return dj;
}

// the small problems' one launch on stream ls (events: its timing)
static void small_launch(bemhip_ctx* ctx, hipStream_t ls, double* G64, const double* d_hphi, double* d_psi,
                         bool events = true) {
    const int nsmall = (int)ctx->small_jobs_h.size();
    SmallJob* dj = small_upload(ctx, ls);
    if (events) BH_HIP(hipEventRecord(ctx->small_ev[0], ls));
    small_solve_launch(ls, dj, nsmall, G64, d_hphi, d_psi, ctx->small_info.as<int>());
    if (events) BH_HIP(hipEventRecord(ctx->small_ev[1], ls));
}

// the small problems' info words and launch time
static void small_finish(bemhip_ctx* ctx, hipStream_t ls, bemhip_solve_stats& sum, bool events = true) {
    const int nsmall = (int)ctx->small_jobs_h.size();
    std::vector<int> info(nsmall);
    BH_HIP(hipMemcpyAsync(info.data(), ctx->small_info.p, sizeof(int) * nsmall, hipMemcpyDeviceToHost, ls));
    BH_HIP(hipStreamSynchronize(ls));
    if (events) {
        float ms = 0.0f;
        BH_HIP(hipEventElapsedTime(&ms, ctx->small_ev[0], ctx->small_ev[1]));
        sum.ms_factor += ms;
    }
    for (int q = 0; q < nsmall; ++q)
        BH_CHECK(info[q] == 0, BEMHIP_ERR_SINGULAR,
                 "zero pivot in column " + std::to_string(info[q]) + " of a batch problem with " +
                     std::to_string(ctx->small_jobs_h[q].n) + " unknowns");
}

// small problems (LU solvers, at most small_n unknowns) into ctx->small_jobs_h,
// the others' indices into big
static void split_small(bemhip_ctx* ctx, const DevTopology& topo, int solver, std::vector<int>& big) {
    big.clear();
    ctx->small_jobs_h.clear();
    for (int q = 0; q < (int)topo.probs.size(); ++q) {
        const AsmProb& pr = topo.probs[q];
        if (solver != BEMHIP_SOLVE_BICGSTAB64 && pr.n <= std::min(ctx->small_n, SMALL_MAX_N))
            ctx->small_jobs_h.push_back(SmallJob{pr.gofs, pr.n, pr.v0});
        else
            big.push_back(q);
    }
    if (!ctx->small_jobs_h.empty() && !ctx->small_ev[0])
        for (auto& e : ctx->small_ev) BH_HIP(hipEventCreate(&e));
}

// lu16_ir64 batches: the problems of each size class factored in lock step
// (class_lu16_enqueue), every (sub-)class on a solver lane of its own (largest
// first; lanes reused round robin when there are more classes), the small
// problems' launch on the next lane.  All work is enqueued before the first
// read-back, so the classes overlap on the device.
struct ClassPlan {
    std::vector<ClassJob> cls;
    std::vector<Solver*> lane;
    int l_small = 0;
    bool small = false;
};

// the plan, with every lane, workspace and device table it needs made ready
// (host work only: none of it is enqueued with the plan's launches)
static ClassPlan plan_classes(bemhip_ctx* ctx, const DevTopology& topo, int solver) {
    ClassPlan P;
    std::vector<int> big;
    split_small(ctx, topo, solver, big);
    std::vector<ClassJob>& cls = P.cls;
    {
        std::vector<int> order(big);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return topo.probs[a].n > topo.probs[b].n; });
        for (int q : order) {
            const AsmProb& pr = topo.probs[q];
            if (cls.empty() || cls.back().n != pr.n) {
                cls.emplace_back();
                cls.back().n = pr.n;
                cls.back().gofs0 = pr.gofs;
            }
            ClassJob& c = cls.back();
            // the G layout puts a class's blocks one after the other, in part order
            BH_CHECK(pr.gofs == c.gofs0 + (int64_t)c.cnt * gblock_stride(pr.n), BEMHIP_ERR_ARG,
                     "batched LU: a size class's G blocks are not evenly spaced");
            c.v0.push_back(pr.v0);
            c.cnt += 1;
        }
        // sub-classes: consecutive problems of a class (still evenly spaced)
        const int split = std::max(1, ctx->batch_split);
        std::vector<ClassJob> sub;
        for (const ClassJob& c : cls) {
            const int parts = std::min(split, c.cnt);
            for (int k = 0, q0 = 0; k < parts; ++k) {
                const int q1 = (int)((int64_t)c.cnt * (k + 1) / parts);
                ClassJob sj;
                sj.n = c.n;
                sj.cnt = q1 - q0;
                sj.gofs0 = c.gofs0 + (int64_t)q0 * gblock_stride(c.n);
                sj.v0.assign(c.v0.begin() + q0, c.v0.begin() + q1);
                sj.lookahead = ctx->batch_lookahead;
                sub.push_back(std::move(sj));
                q0 = q1;
            }
        }
        cls.swap(sub);
    }
    P.small = !ctx->small_jobs_h.empty();
    const int nc = (int)cls.size();
    const int L = std::max(1, std::min(ctx->max_lanes, nc + (P.small ? 1 : 0)));
    while ((int)ctx->lanes.size() < L - 1) {
        ctx->lanes.emplace_back(new Solver());
        ctx->lanes.back()->init(ctx);
    }
    P.lane.resize(L);
    P.lane[0] = ctx;
    for (int l = 1; l < L; ++l) P.lane[l] = ctx->lanes[l - 1].get();
    while ((int)ctx->class_work.size() < nc) ctx->class_work.emplace_back(new ClassWork());
    for (int i = 0; i < nc; ++i) class_lu16_prepare(P.lane[i % L], ctx->class_work[i].get(), cls[i]);
    P.l_small = nc % L;
    if (P.small) small_upload(ctx, P.lane[P.l_small]->stream);
    return P;
}

// every class's factorisation, solves and first sweeps and the small
// problems' launch, forked from and joined back to the context's stream
static void enqueue_classes(bemhip_ctx* ctx, const ClassPlan& P, double* G64, float* G32, const double* d_hphi,
                            double* d_psi, double tol, int maxit, bool events) {
    const int L = (int)P.lane.size(), nc = (int)P.cls.size();
    hipEvent_t ready = ctx->ev[0];
    BH_HIP(hipEventRecord(ready, ctx->stream));
    for (int l = 1; l < L; ++l) BH_HIP(hipStreamWaitEvent(P.lane[l]->stream, ready, 0));
    const int sweeps = std::max(0, std::min(ctx->batch_sweeps, maxit > 0 ? maxit : ctx->batch_sweeps));
    for (int i = 0; i < nc; ++i)
        class_lu16_enqueue(P.lane[i % L], ctx->class_work[i].get(), P.cls[i], G64, G32, d_hphi, d_psi, tol, sweeps,
                           events);
    if (P.small) small_launch(ctx, P.lane[P.l_small]->stream, G64, d_hphi, d_psi, events);
    for (int l = 1; l < L; ++l) {
        // each lane's own join event: a graph capture needs the fork and join
        // of every forked stream
        hipEvent_t done = P.lane[l]->ev_la[3];
        BH_HIP(hipEventRecord(done, P.lane[l]->stream));
        BH_HIP(hipStreamWaitEvent(ctx->stream, done, 0));
    }
}

// read-back, further sweeps, the flagged problems solved again, the small
// problems' info words
static void finish_classes(bemhip_ctx* ctx, const ClassPlan& P, double* G64, float* G32, const double* d_hphi,
                           double* d_psi, double tol, int maxit, bemhip_solve_stats* st, bool events) {
    const int L = (int)P.lane.size(), nc = (int)P.cls.size();
    bemhip_solve_stats sum{};
    sum.converged = 1;
    for (int i = 0; i < nc; ++i) {
        bemhip_solve_stats one{};
        class_lu16_finish(P.lane[i % L], ctx->class_work[i].get(), P.cls[i], G64, G32, d_hphi, d_psi, tol, maxit,
                          &one);
        merge_stats(sum, one);
    }
    if (P.small) small_finish(ctx, P.lane[P.l_small]->stream, sum, events);
    // the context's stream continues after the lanes' further work
    for (int l = 1; l < L; ++l) {
        hipEvent_t done = P.lane[l]->ev_la[3];
        BH_HIP(hipEventRecord(done, P.lane[l]->stream));
        BH_HIP(hipStreamWaitEvent(ctx->stream, done, 0));
    }
    *st = sum;
}

// Independent problems of a batch: the largest first, dealt to up to
// max_lanes solvers by accumulated n^3 (the LU cost); every solver has its
// own streams and work buffers and is driven by its own host thread, so the
// many short, latency-bound kernels of different problems overlap on the
// device.  The solvers start after the assembly (an event on the context's
// stream) and the context's stream waits for all of them.
void solve_batch(bemhip_ctx* ctx, const DevTopology& topo, double* G64, float* G32, const double* d_hphi,
                 double* d_psi, int solver, double tol, int maxit, bemhip_solve_stats* st, bool report_residual) {
    // small problems: one launch for all of them (FP64 GEPP in place, no
    // refinement needed); the rest go to the solver lanes
    std::vector<int> big;
    split_small(ctx, topo, solver, big);
    const int nsmall = (int)ctx->small_jobs_h.size();
    const int np = (int)big.size();
    const int L = std::max(1, std::min(ctx->max_lanes, np));
    while ((int)ctx->lanes.size() < L - 1) {
        ctx->lanes.emplace_back(new Solver());
        ctx->lanes.back()->init(ctx);
    }
    std::vector<Solver*> lane(L);
    lane[0] = ctx;
    for (int l = 1; l < L; ++l) lane[l] = ctx->lanes[l - 1].get();
    std::vector<int> order(big);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return topo.probs[a].n > topo.probs[b].n; });
    std::vector<std::vector<int>> work(L);
    std::vector<double> load(L, 0.0);
    for (int q : order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        work[l].push_back(q);
        load[l] += std::pow((double)topo.probs[q].n, 3.0);
    }
    // the small problems' launch goes last on the least loaded lane: it then
    // fills compute units that the other lanes' last factorisations leave idle
    // (launched first, its work-groups, one per compute unit, held units the
    // large problems' panels were waiting for: 843 against 820 ms per stage
    // for the 256-bubble sweep)
    const int l_small = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    hipEvent_t ready = ctx->ev[0];
    BH_HIP(hipEventRecord(ready, ctx->stream));
    for (int l = 1; l < L; ++l) BH_HIP(hipStreamWaitEvent(lane[l]->stream, ready, 0));
    std::vector<bemhip_solve_stats> part(L);
    std::vector<std::exception_ptr> fail(L);
    auto run = [&](int l) {
        try {
            BH_HIP(hipSetDevice(ctx->device));
            bemhip_solve_stats& sum = part[l];
            sum = bemhip_solve_stats{};
            sum.converged = 1;
            for (int q : work[l]) {
                const AsmProb& pr = topo.probs[q];
                bemhip_solve_stats one{};
                solve_device(lane[l], pr.n, G64 + pr.gofs, G32 ? G32 + pr.gofs : nullptr, d_hphi + pr.v0,
                             d_psi + pr.v0, solver, tol, maxit, &one, report_residual);
                merge_stats(sum, one);
            }
            if (l == l_small && nsmall > 0) small_launch(ctx, lane[l]->stream, G64, d_hphi, d_psi);
        } catch (...) {
            fail[l] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < L; ++l) th.emplace_back(run, l);
    run(0);
    for (auto& t : th) t.join();
    for (int l = 1; l < L; ++l) {
        hipEvent_t done = ctx->ev[1];
        BH_HIP(hipEventRecord(done, lane[l]->stream));
        BH_HIP(hipStreamWaitEvent(ctx->stream, done, 0));
    }
    bemhip_solve_stats sum{};
    sum.converged = 1;
    if (nsmall > 0 && !fail[l_small]) small_finish(ctx, lane[l_small]->stream, sum);
    for (auto& f : fail) if (f) std::rethrow_exception(f);
    for (const auto& p : part) merge_stats(sum, p);
    *st = sum;
}

// An lu16_ir64 batch's assembly and class solves.  With batch_graph the
// device work of a stage (assembly, every class's factorisation, solves and
// first sweeps, the small problems' launch) is captured into a graph the
// second time the same key (topology, buffers, solver settings) comes by, and
// replayed from then on; the read-back and whatever it decides (further
// sweeps, problems solved again) run after it.  A graph is dropped when any
// device buffer has been reallocated since its capture.
static void psi_classes(bemhip_ctx* ctx, const DevTopology& topo, const double* d_xyz, const double* d_phi,
                        uint32_t flags, double tol, int maxit, double* G64, float* G32, double* d_hphi,
                        double* d_psi, bemhip_solve_stats* st) {
    const ClassPlan P = plan_classes(ctx, topo, BEMHIP_SOLVE_LU16_IR64);
    hipEvent_t e0 = ctx->ev[2], e1 = ctx->ev[3];
    BatchGraphKey key{topo.gen, d_xyz, d_phi, d_hphi, d_psi, G64, G32, flags, tol, maxit,
                      ctx->nb16, ctx->batch_split, ctx->batch_lookahead ? 1 : 0, ctx->batch_sweeps, ctx->mm_reserve,
                      ctx->small_n};
    BatchGraph* g = nullptr;
    for (auto& e : ctx->batch_graphs)
        if (e.key == key) g = &e;
    const bool use_graph = ctx->batch_graph && g != nullptr;
    if (use_graph && (!g->exec || g->allocs != g_devbuf_allocs.load())) {
        if (g->exec) {
            (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        BH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        try {
            assemble_device(ctx, topo, d_xyz, d_phi, flags, AsmOut{G64, G32, nullptr, d_hphi});
            enqueue_classes(ctx, P, G64, G32, d_hphi, d_psi, tol, maxit, false);
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        BH_HIP(hipStreamEndCapture(ctx->stream, &graph));
        const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BH_HIP(ie);
        g->allocs = g_devbuf_allocs.load();
        g->captures += 1;
    }
    BH_HIP(hipEventRecord(e0, ctx->stream));
    if (use_graph) {
        BH_HIP(hipGraphLaunch(g->exec, ctx->stream));
        g->launches += 1;
        // the graph ran on the context's stream only: the lanes, on which the
        // read-back and any further sweeps of their classes run, wait for it
        hipEvent_t done = ctx->ev[1];
        BH_HIP(hipEventRecord(done, ctx->stream));
        for (size_t l = 1; l < P.lane.size(); ++l) BH_HIP(hipStreamWaitEvent(P.lane[l]->stream, done, 0));
    } else {
        assemble_device(ctx, topo, d_xyz, d_phi, flags, AsmOut{G64, G32, nullptr, d_hphi});
        BH_HIP(hipEventRecord(e1, ctx->stream));
        enqueue_classes(ctx, P, G64, G32, d_hphi, d_psi, tol, maxit, true);
        if (ctx->batch_graph) {
            // seen once: captured on the next call with this key
            if (ctx->batch_graphs.size() >= 8) {
                if (ctx->batch_graphs.front().exec) (void)hipGraphExecDestroy(ctx->batch_graphs.front().exec);
                ctx->batch_graphs.erase(ctx->batch_graphs.begin());
            }
            ctx->batch_graphs.push_back(BatchGraph{key});
        }
    }
    finish_classes(ctx, P, G64, G32, d_hphi, d_psi, tol, maxit, st, !use_graph);
    if (use_graph) {
        // the replay's span (assembly, factorisations, solves, first sweeps)
        BH_HIP(hipEventRecord(e1, ctx->stream));
        const float ms = event_ms(e0, e1);
        st->ms_factor = ms;
        st->ms_solve = 0.0f;
        st->ms_assemble = 0.0f;
    } else {
        st->ms_assemble = event_ms(e0, e1);
    }
}

// assembly + H phi + solve on device-resident positions / potential
void psi_device(bemhip_ctx* ctx, const DevTopology& topo, const double* d_xyz, const double* d_phi, uint32_t flags,
                int solver, double tol, int maxit, double* d_hphi, double* d_psi, bemhip_solve_stats* st,
                bool report_residual = true) {
    for (const AsmProb& pr : topo.probs)
        BH_CHECK(pr.n <= 65536, BEMHIP_ERR_ARG, "dense solve supports at most 65536 unknowns per problem");
    const size_t nn = (size_t)topo.gsize;
    double* G64 = dev<double>(ctx->G64, nn);
    float* G32 = mixed(solver) ? dev<float>(ctx->G32, nn + MM_C_TAIL) : nullptr;
    hipEvent_t e0 = ctx->ev[2], e1 = ctx->ev[3];
    if (topo.probs.size() > 1 && solver == BEMHIP_SOLVE_LU16_IR64 && ctx->batch_lu) {
        psi_classes(ctx, topo, d_xyz, d_phi, flags, tol, maxit, G64, G32, d_hphi, d_psi, st);
        return;
    }
    BH_HIP(hipEventRecord(e0, ctx->stream));
    assemble_device(ctx, topo, d_xyz, d_phi, flags, AsmOut{G64, G32, nullptr, d_hphi});
    BH_HIP(hipEventRecord(e1, ctx->stream));
    const float ams = event_ms(e0, e1);
    if (topo.probs.size() == 1) {
        solve_device(ctx, (int)topo.N, G64, G32, d_hphi, d_psi, solver, tol, maxit, st, report_residual);
    } else {
        solve_batch(ctx, topo, G64, G32, d_hphi, d_psi, solver, tol, maxit, st, report_residual);
    }
    st->ms_assemble = ams;
}

void accumulate(bemhip_solve_stats& a, const bemhip_solve_stats& s) {
    a.iterations += s.iterations;
    a.ms_assemble += s.ms_assemble;
    a.ms_factor += s.ms_factor;
    a.ms_solve += s.ms_solve;
    a.gemm_ms += s.gemm_ms;
    a.gemm_flops += s.gemm_flops;
    a.gemm_launches += s.gemm_launches;
    a.panels += s.panels;
    a.panels_pivoted += s.panels_pivoted;
    a.residual = std::max(a.residual, s.residual);
    a.correction = std::max(a.correction, s.correction);
    const bool first = a.pad_ == 0;   // pad_ marks a non-empty accumulation
    a.converged = first ? s.converged : (a.converged && s.converged);
    a.pad_ = 1;
}

// resolved per-part Bernoulli parameters [V_0, amp, freq, phase] x nparts on the device
void upload_part_physics(bemhip_sim* s) {
    const int np = s->topo.nparts;
    for (auto* v : {&s->part_v0, &s->part_amp, &s->part_freq, &s->part_phase})
        if ((int)v->size() != np) v->assign(np, std::numeric_limits<double>::quiet_NaN());
    std::vector<double> h(4 * (size_t)np);
    auto pick = [](double a, double d) { return std::isnan(a) ? d : a; };
    for (int c = 0; c < np; ++c) {
        h[4 * c + 0] = pick(s->part_v0[c], s->P.V_0);
        h[4 * c + 1] = pick(s->part_amp[c], s->P.field_amp);
        h[4 * c + 2] = pick(s->part_freq[c], s->P.field_freq);
        h[4 * c + 3] = pick(s->part_phase[c], s->P.field_phase);
    }
    double* d = dev<double>(s->pphys, h.size());
    h2d(d, h.data(), sizeof(double) * h.size(), s->ctx->stream);
    BH_HIP(hipStreamSynchronize(s->ctx->stream));
}

// signed volume of every loose part at positions x (host copy)
std::vector<double> part_volumes_host(bemhip_sim* s, const double* x) {
    double* tvol = dev<double>(s->tvol, s->M);
    double* pvol = dev<double>(s->pvol, s->topo.nparts);
    tri_geometry(s->ctx, s->topo, x, nullptr, nullptr, nullptr, tvol);
    part_volumes(s->ctx, s->topo, tvol, pvol);
    std::vector<double> pv(s->topo.nparts);
    d2h(pv.data(), pvol, sizeof(double) * pv.size(), s->ctx->stream);
    return pv;
}

StageParams stage_params(const bemhip_sim* s) {
    StageParams sp;
    sp.p_inf = s->P.p_inf; sp.epsilon = s->P.epsilon; sp.sigma = s->P.sigma; sp.gamma = s->P.gamma;
    return sp;
}

// LinLinSim::position_t (LinLinSim.cpp:26-83) followed by pot_t_multi
// (LinLinSim.cpp:98-120) for one stage: positions xs, potential ps ->
// velocities kx, Dphi/Dt kp
void stage(bemhip_sim* s, const double* xs, const double* ps, double t, double* kx, double* kp) {
    bemhip_ctx* ctx = s->ctx;
    const int64_t N = s->N, M = s->M;
    double* hphi = dev<double>(s->hphi, N);
    double* psil = dev<double>(s->psil, N);
    psi_device(ctx, s->topo, xs, ps, s->P.flags, s->P.solver, s->P.tol, s->P.maxit, hphi, psil, &s->last, false);
    accumulate(s->accum, s->last);
    double* tnorm = dev<double>(s->tnorm, 3 * M);
    double* tgrad = dev<double>(s->tgrad, 3 * M);
    double* tvol = dev<double>(s->tvol, M);
    tri_geometry(ctx, s->topo, xs, ps, tnorm, tgrad, tvol);
    vertex_velocity(ctx, s->topo, tnorm, tgrad, psil, kx);
    if (!kp) return;
    double* pvol = dev<double>(s->pvol, s->topo.nparts);
    part_volumes(ctx, s->topo, tvol, pvol);
    double* tk = nullptr;
    if (s->P.sigma != 0.0) {
        double* vn = dev<double>(s->vn, 3 * N);
        tk = dev<double>(s->tk, M);
        vertex_normals(ctx, s->topo, xs, vn);
        tri_curvature(ctx, s->topo, xs, vn, tk);
    }
    vertex_dpot(ctx, s->topo, stage_params(s), xs, kx, tk, pvol, s->pphys.as<double>(), t, kp);
}

// ColocSimPin (ColocSimPin.cpp:21-118) + solve_system: the unknowns are psi at
// the free vertices and phi at the pinned ones; pot holds phi at the free
// vertices and the prescribed psi at the pinned ones.  H pot is formed in the
// assembly over the free columns (pinned entries zeroed), and the column swap
// adds the pinned columns' part.
void psi_pinned(bemhip_sim* s, const double* xs, const double* pot, double* rhs, double* x,
                bemhip_solve_stats* st, bool report_residual = true) {
    bemhip_ctx* ctx = s->ctx;
    const int64_t N = s->N, np = s->n_pin;
    BH_CHECK(N <= 65536, BEMHIP_ERR_ARG, "dense solve supports at most 65536 unknowns per problem");
    const size_t nn = (size_t)N * N;
    double* G64 = dev<double>(ctx->G64, nn);
    float* G32 = mixed(s->P.solver) ? dev<float>(ctx->G32, nn + MM_C_TAIL) : nullptr;
    double* H = dev<double>(ctx->H, nn);
    double* pz = dev<double>(s->pz, N);
    pin_zero_tail(ctx, N, np, pot, pz);
    hipEvent_t e0 = ctx->ev[2], e1 = ctx->ev[3];
    BH_HIP(hipEventRecord(e0, ctx->stream));
    assemble_device(ctx, s->topo, xs, pz, s->P.flags, AsmOut{G64, G32, H, rhs});
    pin_swap_columns(ctx, N, np, G64, G32, H, pot, rhs);
    BH_HIP(hipEventRecord(e1, ctx->stream));
    const float ams = event_ms(e0, e1);
    solve_device(ctx, (int)N, G64, G32, rhs, x, s->P.solver, s->P.tol, s->P.maxit, st, report_residual);
    st->ms_assemble = ams;
}

// ColocSimPin::position_t (ColocSimPin.cpp:139-194) + pot_t_multi.  As in the
// reference, ps is rewritten: its pinned entries take the state's psi for the
// solve and the solved phi afterwards.
void stage_pinned(bemhip_sim* s, const double* xs, double* ps, double t, double* kx, double* kp) {
    bemhip_ctx* ctx = s->ctx;
    const int64_t N = s->N, M = s->M;
    double* tnorm = dev<double>(s->tnorm, 3 * M);
    double* tgrad = dev<double>(s->tgrad, 3 * M);
    double* tvol = dev<double>(s->tvol, M);
    tri_geometry(ctx, s->topo, xs, ps, tnorm, tgrad, tvol);   // tangential gradients of the stage phi first
    pin_copy_tail(ctx, N, s->n_pin, s->psi.as<double>(), ps);
    double* hphi = dev<double>(s->hphi, N);
    double* psil = dev<double>(s->psil, N);
    psi_pinned(s, xs, ps, hphi, psil, &s->last, false);
    accumulate(s->accum, s->last);
    vertex_velocity(ctx, s->topo, tnorm, tgrad, psil, kx);
    double* vn = dev<double>(s->vn, 3 * N);
    vertex_normals(ctx, s->topo, xs, vn);
    pin_drift(ctx, N, s->n_pin, vn, psil, kx, ps);
    if (!kp) return;
    double* pvol = dev<double>(s->pvol, s->topo.nparts);
    part_volumes(ctx, s->topo, tvol, pvol);
    double* tk = nullptr;
    if (s->P.sigma != 0.0) {
        tk = dev<double>(s->tk, M);
        tri_curvature(ctx, s->topo, xs, vn, tk);
    }
    vertex_dpot(ctx, s->topo, stage_params(s), xs, kx, tk, pvol, s->pphys.as<double>(), t, kp);
}

// a stage of the time steppers: the pinned variant may rewrite ps
void stage_step(bemhip_sim* s, const double* xs, double* ps, double t, double* kx, double* kp) {
    if (s->n_pin > 0) stage_pinned(s, xs, ps, t, kx, kp);
    else stage(s, xs, ps, t, kx, kp);
}

// psi of the state (x, pot) by the simulation's own formulation
void sim_psi(bemhip_sim* s, const double* x, const double* pot, double* rhs, double* psi, bemhip_solve_stats* st) {
    if (s->n_pin > 0) psi_pinned(s, x, pot, rhs, psi, st);
    else psi_device(s->ctx, s->topo, x, pot, s->P.flags, s->P.solver, s->P.tol, s->P.maxit, rhs, psi, st);
}

double get_dt(bemhip_sim* s, double dp, const double* kx, const double* kp) {  // Simulation.cpp:77-92
    double m[2];
    dt_reduce(s->ctx, s->N, kx, kp, m);
    double dt = dp / (m[0] + s->P.dp_balance * std::sqrt(m[1]));
    if (s->P.min_dt > 0.0) dt = std::min(dt, s->P.min_dt);
    return dt;
}

}  // namespace

// =============================================================== C-ABI ==
extern "C" {

int bemhip_create(int device, bemhip_ctx** out) {
    if (!out) return BEMHIP_ERR_ARG;
    *out = nullptr;
    auto ctx = std::make_unique<bemhip_ctx>();
    int rc = guarded(ctx.get(), [&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
            throw Error(BEMHIP_ERR_NO_DEVICE, "no HIP device " + std::to_string(device));
        BH_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        BH_HIP(hipGetDeviceProperties(&prop, device));
        BH_CHECK(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, BEMHIP_ERR_NO_DEVICE,
                 std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        ctx->device = device;
        ctx->cus = prop.multiProcessorCount;
        exterior_scratch_init(ctx.get());
        for (auto& e : ctx->ev) BH_HIP(hipEventCreate(&e));
        if (const char* v = std::getenv("BEMHIP_LA_MIN_REST")) ctx->la_min_rest = std::atoi(v);
        const char* la = std::getenv("BEMHIP_LU_LOOKAHEAD");
        if (la && la[0] == '0') ctx->lookahead = false;
        const char* sp = std::getenv("BEMHIP_LU_SPECULATE");
        if (sp && sp[0] == '0') ctx->speculate = false;
        const char* nt = std::getenv("BEMHIP_NP_TOP");
        if (nt && nt[0] == '0') ctx->np_top = false;
        const char* n32 = std::getenv("BEMHIP_NP_TOP32");
        if (n32 && n32[0] == '0') ctx->np_top32 = false;
        if (const char* v = std::getenv("BEMHIP_NB64")) ctx->nb64 = std::atoi(v);
        if (const char* v = std::getenv("BEMHIP_NB32")) ctx->nb32 = std::atoi(v);
        if (const char* v = std::getenv("BEMHIP_NB_SPLIT")) ctx->nb_split = std::atoi(v);
        if (const char* v = std::getenv("BEMHIP_NB16")) ctx->nb16 = std::atoi(v);
        if (const char* v = std::getenv("BEMHIP_MM_RESERVE")) ctx->mm_reserve = std::max(0, std::atoi(v));
        const char* tp = std::getenv("BEMHIP_TRSV_PERSISTENT");
        if (tp && tp[0] == '0') ctx->trsv_persistent = false;
        const char* tg = std::getenv("BEMHIP_TRSV_FORCE_GIVEUP");
        if (tg && tg[0] == '1') ctx->trsv_force_giveup = true;
        if (const char* v = std::getenv("BEMHIP_LANES")) ctx->max_lanes = std::max(1, std::atoi(v));
        if (const char* v = std::getenv("BEMHIP_SMALL_N")) ctx->small_n = std::max(0, std::atoi(v));
        const char* s1 = std::getenv("BEMHIP_SMALL_SINGLE");
        if (s1 && s1[0] == '1') ctx->small_single = true;
        const char* sf = std::getenv("BEMHIP_SPLIT_FIXED");
        if (sf && sf[0] == '0') ctx->split_fixed = false;
        const char* bl = std::getenv("BEMHIP_BATCH_LU");
        if (bl && bl[0] == '0') ctx->batch_lu = false;
        if (const char* v = std::getenv("BEMHIP_BATCH_SWEEPS")) ctx->batch_sweeps = std::max(0, std::atoi(v));
        if (const char* v = std::getenv("BEMHIP_BATCH_FORCE_REDO")) ctx->batch_force_redo = std::atoi(v);
        if (const char* v = std::getenv("BEMHIP_BATCH_SPLIT")) ctx->batch_split = std::max(1, std::atoi(v));
        const char* lm = std::getenv("BEMHIP_LU16_MERGE");
        if (lm && lm[0] == '0') ctx->lu16_merge = false;
        if (const char* v = std::getenv("BEMHIP_LU16_BULK_FIRST")) ctx->lu16_bulk_first = std::atoll(v);
        const char* bla = std::getenv("BEMHIP_BATCH_LOOKAHEAD");
        if (bla && bla[0] == '0') ctx->batch_lookahead = false;
        const char* bg = std::getenv("BEMHIP_BATCH_GRAPH");
        if (bg && bg[0] == '1') ctx->batch_graph = true;
        ctx->init(nullptr);
    });
    if (rc != BEMHIP_OK) {
        // callers get no handle: release whatever was created
        for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
        ctx->release();
        return rc;
    }
    *out = ctx.release();
    return BEMHIP_OK;
}

int bemhip_destroy(bemhip_ctx* ctx) {
    if (!ctx) return BEMHIP_OK;
    (void)hipSetDevice(ctx->device);
    for (auto& g : ctx->batch_graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->batch_graphs.clear();
    for (auto& l : ctx->lanes) l->release();
    ctx->class_work.clear();
    ctx->lanes.clear();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->small_ev) if (e) (void)hipEventDestroy(e);
    ctx->small_jobs.release();
    ctx->small_info.release();
    ctx->release();
    delete ctx;
    return BEMHIP_OK;
}

const char* bemhip_last_error(const bemhip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void bemhip_quadrature(double* o) {
    for (int q = 0; q < 7; ++q) {
        o[q] = GQ_X[q];
        o[7 + q] = GQ_W[q];
        o[14 + q] = TQ_X[q];
        o[21 + q] = TQ_Y[q];
        o[28 + q] = TQ_W[q];
    }
}

int bemhip_assemble(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri, uint32_t flags,
                    double* G, double* H, const double* phi, double* Hphi) {
    if (!ctx) return BEMHIP_ERR_ARG;
    return guarded(ctx, [&] {
        BH_CHECK(xyz && tri && (G || H || (phi && Hphi)), BEMHIP_ERR_ARG, "null argument");
        BH_CHECK(N <= 46340 || !(G || H), BEMHIP_ERR_ARG, "host G output limited to N <= 46340");
        BH_CHECK(!(flags & ~BEMHIP_MIRROR_X0), BEMHIP_ERR_ARG, "unknown assembly flag");
        check_finite(phi, N, "phi");
        Topology th = build_topology(N, xyz, M, tri);
        DevTopology topo;
        topo.upload(th, ctx->stream);
        double* dx = dev<double>(ctx->xyz, 3 * N);
        h2d(dx, xyz, sizeof(double) * 3 * N, ctx->stream);
        double* dphi = nullptr;
        double* dh = nullptr;
        if (phi && Hphi) {
            dphi = dev<double>(ctx->phi, N);
            h2d(dphi, phi, sizeof(double) * N, ctx->stream);
            dh = dev<double>(ctx->hphi, N);
        }
        const size_t nn = (size_t)N * N;
        double* dG = dev<double>(ctx->G64, nn);
        double* dH = H ? dev<double>(ctx->H, nn) : nullptr;
        assemble_device(ctx, topo, dx, dphi, flags, AsmOut{dG, nullptr, dH, dh});
        if (G) d2h(G, dG, sizeof(double) * nn, ctx->stream);
        if (H) d2h(H, dH, sizeof(double) * nn, ctx->stream);
        if (dh) d2h(Hphi, dh, sizeof(double) * N, ctx->stream);
    });
}

int bemhip_solve(bemhip_ctx* ctx, int64_t N, const double* G, const double* b, double* x, int solver, double tol,
                 int32_t maxit, bemhip_solve_stats* stats) {
    if (!ctx) return BEMHIP_ERR_ARG;
    return guarded(ctx, [&] {
        BH_CHECK(G && b && x && N > 0 && N <= 46340, BEMHIP_ERR_ARG, "bad arguments");
        check_finite(b, N, "b");
        const size_t nn = (size_t)N * N;
        double* dG = dev<double>(ctx->G64, nn);
        h2d(dG, G, sizeof(double) * nn, ctx->stream);
        float* dG32 = nullptr;
        if (mixed(solver)) {
            dG32 = dev<float>(ctx->G32, nn + MM_C_TAIL);
            k_d2f<<<4096, 256, 0, ctx->stream>>>(nn, dG, dG32);
            BH_HIP(hipGetLastError());
        }
        double* db = dev<double>(ctx->vec[10], N);
        double* dx = dev<double>(ctx->vec[11], N);
        h2d(db, b, sizeof(double) * N, ctx->stream);
        bemhip_solve_stats st{};
        solve_device(ctx, (int)N, dG, dG32, db, dx, solver, tol, maxit, &st);
        d2h(x, dx, sizeof(double) * N, ctx->stream);
        if (stats) *stats = st;
        check_finite(x, N, "solution");
        if (!st.converged) throw Error(BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
    });
}

int bemhip_compute_psi(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                       const double* phi, uint32_t flags, int solver, double tol, int32_t maxit, double* psi,
                       bemhip_solve_stats* stats) {
    if (!ctx) return BEMHIP_ERR_ARG;
    return guarded(ctx, [&] {
        BH_CHECK(xyz && tri && phi && psi, BEMHIP_ERR_ARG, "null argument");
        BH_CHECK(!(flags & ~(BEMHIP_MIRROR_X0 | BEMHIP_INDEPENDENT_PARTS)), BEMHIP_ERR_ARG, "unknown flag");
        check_finite(phi, N, "phi");
        Topology th = build_topology(N, xyz, M, tri, (flags & BEMHIP_INDEPENDENT_PARTS) != 0);
        DevTopology topo;
        topo.upload(th, ctx->stream);
        double* dx = dev<double>(ctx->xyz, 3 * N);
        double* dphi = dev<double>(ctx->phi, N);
        double* dh = dev<double>(ctx->hphi, N);
        double* dpsi = dev<double>(ctx->vec[11], N);
        h2d(dx, xyz, sizeof(double) * 3 * N, ctx->stream);
        h2d(dphi, phi, sizeof(double) * N, ctx->stream);
        bemhip_solve_stats st{};
        psi_device(ctx, topo, dx, dphi, flags, solver, tol, maxit, dh, dpsi, &st);
        d2h(psi, dpsi, sizeof(double) * N, ctx->stream);
        if (stats) *stats = st;
        check_finite(psi, N, "psi");
        if (!st.converged) throw Error(BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
    });
}

void bemhip_sim_default_params(bemhip_sim_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->p_inf = 1.0; p->epsilon = 1.0; p->sigma = 0.0; p->gamma = 1.0;   // Simulation.hpp:29
    p->V_0 = std::numeric_limits<double>::quiet_NaN();
    p->dp_balance = 3.0;                                                // Simulation.hpp:37
    p->min_dt = -1.0;                                                   // Simulation.hpp:36
    p->wall_eps = 1e-2;                                                 // LinLinSim.hpp:24
    p->tol = 0.0;
    p->maxit = 0;
    p->solver = BEMHIP_SOLVE_BICGSTAB64;                                // Simulation.hpp:38
    p->flags = 0;
    p->wall_guards = 0;
    p->frozen_stage_time = 1;
}

int bemhip_sim_create(bemhip_ctx* ctx, const bemhip_sim_params* p, int64_t N, const double* xyz, int64_t M,
                      const uint32_t* tri, bemhip_sim** out) {
    if (!ctx || !out) return BEMHIP_ERR_ARG;
    *out = nullptr;
    auto s = std::make_unique<bemhip_sim>();
    int rc = guarded(ctx, [&] {
        BH_CHECK(p && xyz && tri, BEMHIP_ERR_ARG, "null argument");
        BH_CHECK(!(p->flags & ~(BEMHIP_MIRROR_X0 | BEMHIP_INDEPENDENT_PARTS)), BEMHIP_ERR_ARG, "unknown flag");
        BH_CHECK(p->solver >= 0 && p->solver <= BEMHIP_SOLVE_LU16_IR64, BEMHIP_ERR_ARG, "unknown solver");
        s->ctx = ctx;
        s->P = *p;
        s->N = N;
        s->M = M;
        s->th = build_topology(N, xyz, M, tri, (p->flags & BEMHIP_INDEPENDENT_PARTS) != 0);
        s->topo.upload(s->th, ctx->stream);
        double* x = dev<double>(s->x, 3 * N);
        h2d(x, xyz, sizeof(double) * 3 * N, ctx->stream);
        BH_HIP(hipMemsetAsync(dev<double>(s->phi, N), 0, sizeof(double) * N, ctx->stream));   // LinLinSim.hpp:28-29
        BH_HIP(hipMemsetAsync(dev<double>(s->psi, N), 0, sizeof(double) * N, ctx->stream));
        if (std::isnan(s->P.V_0)) {   // Simulation.hpp:42: total volume of the initial mesh
            const std::vector<double> pv = part_volumes_host(s.get(), x);
            double v = 0.0;
            for (double a : pv) v += a;
            s->P.V_0 = v;
            // a batch of independent bubbles: each part is its own simulation
            if (s->th.independent) s->part_v0 = pv;
        }
        upload_part_physics(s.get());
    });
    if (rc != BEMHIP_OK) return rc;
    *out = s.release();
    return BEMHIP_OK;
}

int bemhip_sim_create_pinned(bemhip_ctx* ctx, const bemhip_sim_params* p, int64_t N, const double* xyz, int64_t M,
                             const uint32_t* tri, const double* phi, const double* psi, bemhip_sim** out,
                             int64_t* npin) {
    if (!ctx || !out || !p || !xyz || !tri || N <= 0 || M <= 0) return BEMHIP_ERR_ARG;
    *out = nullptr;
    if (p->flags & BEMHIP_INDEPENDENT_PARTS) {
        ctx->err = "pinned vertices need the coupled problem";
        return BEMHIP_ERR_ARG;
    }
    // Simulation(initial): V_0 from the mesh as given (Simulation.hpp:42)
    bemhip_sim* s = nullptr;
    int rc = bemhip_sim_create(ctx, p, N, xyz, M, tri, &s);
    if (rc != BEMHIP_OK) return rc;
    rc = guarded(ctx, [&] {
        std::vector<double> x(xyz, xyz + 3 * N), ph(N, 0.0), ps(N, 0.0);
        std::vector<uint32_t> t(tri, tri + 3 * M);
        if (phi) { check_finite(phi, N, "phi"); ph.assign(phi, phi + N); }
        if (psi) { check_finite(psi, N, "psi"); ps.assign(psi, psi + N); }
        int64_t np = 0;
        const int r = bemhip_pin_boundary(N, x.data(), M, t.data(), ph.data(), ps.data(), &np);
        BH_CHECK(r == BEMHIP_OK, r, "bad mesh");
        const double time = s->time;
        const int r2 = bemhip_sim_set_mesh(s, N, x.data(), M, t.data());
        BH_CHECK(r2 == BEMHIP_OK, r2, ctx->err);
        h2d(s->phi.as<double>(), ph.data(), sizeof(double) * N, ctx->stream);
        h2d(s->psi.as<double>(), ps.data(), sizeof(double) * N, ctx->stream);
        BH_HIP(hipStreamSynchronize(ctx->stream));   // the host vectors end with this scope
        s->time = time;
        s->n_pin = np;
        if (npin) *npin = np;
    });
    if (rc != BEMHIP_OK) {
        bemhip_sim_destroy(s);
        return rc;
    }
    *out = s;
    return BEMHIP_OK;
}

int bemhip_sim_set_pinned(bemhip_sim* s, int64_t npin) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(npin >= 0 && npin <= s->N, BEMHIP_ERR_ARG, "pinned count out of range");
        BH_CHECK(npin == 0 || !(s->P.flags & BEMHIP_INDEPENDENT_PARTS), BEMHIP_ERR_ARG,
                 "pinned vertices need the coupled problem");
        s->n_pin = npin;
    });
}

int bemhip_sim_get_pinned(bemhip_sim* s, int64_t* npin) {
    if (!s || !npin) return BEMHIP_ERR_ARG;
    *npin = s->n_pin;
    return BEMHIP_OK;
}

int bemhip_sim_destroy(bemhip_sim* s) {
    if (!s) return BEMHIP_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return BEMHIP_OK;
}

int bemhip_sim_set_phi(bemhip_sim* s, const double* phi) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(phi, BEMHIP_ERR_ARG, "null phi");
        check_finite(phi, s->N, "phi");
        h2d(s->phi.p, phi, sizeof(double) * s->N, s->ctx->stream);
        BH_HIP(hipStreamSynchronize(s->ctx->stream));
    });
}

int bemhip_sim_set_psi(bemhip_sim* s, const double* psi) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(psi, BEMHIP_ERR_ARG, "null psi");
        h2d(s->psi.p, psi, sizeof(double) * s->N, s->ctx->stream);
        BH_HIP(hipStreamSynchronize(s->ctx->stream));
    });
}

int bemhip_sim_set_V_0(bemhip_sim* s, double v) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        s->P.V_0 = v;
        s->part_v0.assign(s->topo.nparts, std::numeric_limits<double>::quiet_NaN());
        upload_part_physics(s);
    });
}

int bemhip_sim_set_part_physics(bemhip_sim* s, int32_t nparts, const double* V_0, const double* amp,
                                const double* freq, const double* phase) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(nparts == s->topo.nparts, BEMHIP_ERR_ARG,
                 "mesh has " + std::to_string(s->topo.nparts) + " parts, got " + std::to_string(nparts));
        auto set = [&](std::vector<double>& dst, const double* src) {
            if (!src) return;   // unchanged
            check_finite(src, nparts, "per-part parameter");
            dst.assign(src, src + nparts);
        };
        set(s->part_v0, V_0);
        set(s->part_amp, amp);
        set(s->part_freq, freq);
        set(s->part_phase, phase);
        upload_part_physics(s);
    });
}

int bemhip_sim_get_parts(bemhip_sim* s, int32_t* nparts) {
    if (!s || !nparts) return BEMHIP_ERR_ARG;
    *nparts = s->topo.nparts;
    return BEMHIP_OK;
}

int bemhip_sim_get_part_volumes(bemhip_sim* s, double* vol, int32_t nparts) {
    if (!s || !vol) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(nparts == s->topo.nparts, BEMHIP_ERR_ARG,
                 "mesh has " + std::to_string(s->topo.nparts) + " parts, got " + std::to_string(nparts));
        const std::vector<double> pv = part_volumes_host(s, s->x.as<double>());
        std::copy(pv.begin(), pv.end(), vol);
    });
}

int bemhip_sim_set_time(bemhip_sim* s, double t) {
    if (!s) return BEMHIP_ERR_ARG;
    s->time = t;
    return BEMHIP_OK;
}

int bemhip_sim_set_solver(bemhip_sim* s, int solver, double tol, int32_t maxit) {
    if (!s) return BEMHIP_ERR_ARG;
    if (solver < 0 || solver > BEMHIP_SOLVE_LU16_IR64) { s->ctx->err = "unknown solver"; return BEMHIP_ERR_ARG; }
    s->P.solver = solver;
    s->P.tol = tol;
    s->P.maxit = maxit;
    return BEMHIP_OK;
}

int bemhip_sim_set_params(bemhip_sim* s, const bemhip_sim_params* p) {
    if (!s || !p) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(!(p->flags & ~(BEMHIP_MIRROR_X0 | BEMHIP_INDEPENDENT_PARTS)), BEMHIP_ERR_ARG, "unknown flag");
        BH_CHECK(!((p->flags ^ s->P.flags) & BEMHIP_INDEPENDENT_PARTS), BEMHIP_ERR_ARG,
                 "BEMHIP_INDEPENDENT_PARTS is fixed when the simulation is created");
        BH_CHECK(p->solver >= 0 && p->solver <= BEMHIP_SOLVE_LU16_IR64, BEMHIP_ERR_ARG, "unknown solver");
        const double v0 = s->P.V_0;
        s->P = *p;
        if (std::isnan(p->V_0)) s->P.V_0 = v0;
        else s->part_v0.assign(s->topo.nparts, std::numeric_limits<double>::quiet_NaN());
        upload_part_physics(s);
    });
}

int bemhip_sim_set_mesh(bemhip_sim* s, int64_t N, const double* xyz, int64_t M, const uint32_t* tri) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(xyz && tri, BEMHIP_ERR_ARG, "null argument");
        bemhip_ctx* ctx = s->ctx;
        BH_CHECK(s->n_pin <= N, BEMHIP_ERR_ARG, "the new mesh has fewer vertices than are pinned");
        Topology th = build_topology(N, xyz, M, tri, (s->P.flags & BEMHIP_INDEPENDENT_PARTS) != 0);
        const int old_parts = s->topo.nparts;
        s->th = std::move(th);
        s->topo.upload(s->th, ctx->stream);
        s->N = N;
        s->M = M;
        double* x = dev<double>(s->x, 3 * N);
        h2d(x, xyz, sizeof(double) * 3 * N, ctx->stream);
        BH_HIP(hipMemsetAsync(dev<double>(s->phi, N), 0, sizeof(double) * N, ctx->stream));
        BH_HIP(hipMemsetAsync(dev<double>(s->psi, N), 0, sizeof(double) * N, ctx->stream));
        if (s->topo.nparts != old_parts) {
            // per-part parameters no longer match: back to the scalars, and in a
            // batch every part starts from its own volume
            for (auto* v : {&s->part_v0, &s->part_amp, &s->part_freq, &s->part_phase}) v->clear();
            if (s->th.independent) s->part_v0 = part_volumes_host(s, x);
        }
        upload_part_physics(s);
    });
}

int bemhip_sim_get_state(bemhip_sim* s, double* xyz, double* phi, double* psi, double* time) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        if (xyz) d2h(xyz, s->x.p, sizeof(double) * 3 * s->N, s->ctx->stream);
        if (phi) d2h(phi, s->phi.p, sizeof(double) * s->N, s->ctx->stream);
        if (psi) d2h(psi, s->psi.p, sizeof(double) * s->N, s->ctx->stream);
        if (time) *time = s->time;
    });
}

int bemhip_sim_get_triangles(bemhip_sim* s, uint32_t* tri) {
    if (!s || !tri) return BEMHIP_ERR_ARG;
    std::copy(s->th.tri.begin(), s->th.tri.end(), tri);
    return BEMHIP_OK;
}

int bemhip_sim_export_ply(bemhip_sim* s, const char* path, int32_t precision) {
    if (!s || !path) return BEMHIP_ERR_ARG;
    std::vector<double> x(3 * s->N), ph(s->N), ps(s->N);
    int rc = bemhip_sim_get_state(s, x.data(), ph.data(), ps.data(), nullptr);
    if (rc != BEMHIP_OK) return rc;
    rc = bemhip_ply_write(path, s->N, x.data(), s->M, s->th.tri.data(), ph.data(), ps.data(), precision);
    if (rc != BEMHIP_OK) s->ctx->err = bemhip_io_last_error();
    return rc;
}

int bemhip_sim_get_volume(bemhip_sim* s, double* vol) {
    if (!s || !vol) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        const std::vector<double> pv = part_volumes_host(s, s->x.as<double>());
        double v = 0.0;
        for (double a : pv) v += a;
        *vol = v;
    });
}

int bemhip_sim_position_t(bemhip_sim* s, double* vel, double* psi) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        double* kx = dev<double>(s->kx, 12 * s->N);
        // on a copy of phi: the pinned stage rewrites its potential
        double* ps = dev<double>(s->ps, s->N);
        BH_HIP(hipMemcpyAsync(ps, s->phi.p, sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->ctx->stream));
        stage_step(s, s->x.as<double>(), ps, s->time, kx, nullptr);
        if (vel) d2h(vel, kx, sizeof(double) * 3 * s->N, s->ctx->stream);
        if (psi) d2h(psi, s->psil.p, sizeof(double) * s->N, s->ctx->stream);
        BH_CHECK(s->last.converged, BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
    });
}

int bemhip_sim_pot_t_multi(bemhip_sim* s, const double* vel, double t, double* dpot) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(vel && dpot, BEMHIP_ERR_ARG, "null argument");
        const int64_t N = s->N, M = s->M;
        double* kx = dev<double>(s->kx, 12 * N);
        double* kp = dev<double>(s->kp, 4 * N);
        h2d(kx, vel, sizeof(double) * 3 * N, s->ctx->stream);
        double* tvol = dev<double>(s->tvol, M);
        double* pvol = dev<double>(s->pvol, s->topo.nparts);
        tri_geometry(s->ctx, s->topo, s->x.as<double>(), nullptr, nullptr, nullptr, tvol);
        part_volumes(s->ctx, s->topo, tvol, pvol);
        double* tk = nullptr;
        if (s->P.sigma != 0.0) {
            double* vn = dev<double>(s->vn, 3 * N);
            tk = dev<double>(s->tk, M);
            vertex_normals(s->ctx, s->topo, s->x.as<double>(), vn);
            tri_curvature(s->ctx, s->topo, s->x.as<double>(), vn, tk);
        }
        vertex_dpot(s->ctx, s->topo, stage_params(s), s->x.as<double>(), kx, tk, pvol, s->pphys.as<double>(), t,
                    kp);
        d2h(dpot, kp, sizeof(double) * N, s->ctx->stream);
    });
}

int bemhip_sim_compute_psi(bemhip_sim* s) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        double* hphi = dev<double>(s->hphi, s->N);
        sim_psi(s, s->x.as<double>(), s->phi.as<double>(), hphi, s->psi.as<double>(), &s->last);
        accumulate(s->accum, s->last);
        BH_HIP(hipStreamSynchronize(s->ctx->stream));
        BH_CHECK(s->last.converged, BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
    });
}

int bemhip_exterior_potential(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                              const double* phi, const double* psi, int64_t P, const double* points,
                              double* out) {
    if (!ctx) return BEMHIP_ERR_ARG;
    return guarded(ctx, [&] {
        BH_CHECK(N > 0 && M > 0 && P >= 0, BEMHIP_ERR_ARG, "bad sizes");
        BH_CHECK(xyz && tri && phi && psi && (P == 0 || (points && out)), BEMHIP_ERR_ARG, "null argument");
        for (int64_t i = 0; i < 3 * M; ++i) BH_CHECK(tri[i] < (uint64_t)N, BEMHIP_ERR_BAD_MESH, "vertex index out of range");
        check_finite(phi, N, "phi");
        check_finite(psi, N, "psi");
        check_finite(points, 3 * P, "points");
        if (P == 0) return;
        double* dx = dev<double>(ctx->xyz, 3 * N);
        double* dphi = dev<double>(ctx->phi, N);
        double* dpsi = dev<double>(ctx->ext_psi, N);
        uint32_t* dtri = dev<uint32_t>(ctx->ext_tri, 3 * M);
        double* dtg = dev<double>(ctx->ext_tgt, 3 * P);
        double* dout = dev<double>(ctx->ext_out, P);
        h2d(dx, xyz, sizeof(double) * 3 * N, ctx->stream);
        h2d(dphi, phi, sizeof(double) * N, ctx->stream);
        h2d(dpsi, psi, sizeof(double) * N, ctx->stream);
        h2d(dtri, tri, sizeof(uint32_t) * 3 * M, ctx->stream);
        h2d(dtg, points, sizeof(double) * 3 * P, ctx->stream);
        exterior_device(ctx, M, dx, dtri, dphi, dpsi, P, dtg, dout);
        d2h(out, dout, sizeof(double) * P, ctx->stream);
    });
}

int bemhip_exterior_potential_noalloc(bemhip_ctx* ctx, int64_t N, const double* d_xyz, int64_t M,
                                      const uint32_t* d_tri, const double* d_phi, const double* d_psi, int64_t P,
                                      const double* d_points, double* d_out) {
    if (!ctx) return BEMHIP_ERR_ARG;
    return guarded(ctx, [&] {
        BH_CHECK(N > 0 && M > 0 && P >= 0, BEMHIP_ERR_ARG, "bad sizes");
        BH_CHECK(d_xyz && d_tri && d_phi && d_psi && (P == 0 || (d_points && d_out)), BEMHIP_ERR_ARG,
                 "null argument");
        if (P == 0) return;
        exterior_device_noalloc(ctx, N, M, d_xyz, d_tri, d_phi, d_psi, P, d_points, d_out);
    });
}

int bemhip_sim_exterior_pot(bemhip_sim* s, int64_t P, const double* points, double* out) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(P >= 0 && (P == 0 || (points && out)), BEMHIP_ERR_ARG, "bad arguments");
        check_finite(points, 3 * P, "points");
        if (P == 0) return;
        bemhip_ctx* ctx = s->ctx;
        const int64_t N = s->N;
        // psi for the current state into scratch, the simulation's psi untouched
        double* hphi = dev<double>(s->hphi, N);
        double* psi = dev<double>(ctx->ext_psi, N);
        bemhip_solve_stats st{};
        sim_psi(s, s->x.as<double>(), s->phi.as<double>(), hphi, psi, &st);
        BH_CHECK(st.converged, BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
        double* dtg = dev<double>(ctx->ext_tgt, 3 * P);
        double* dout = dev<double>(ctx->ext_out, P);
        h2d(dtg, points, sizeof(double) * 3 * P, ctx->stream);
        exterior_device(ctx, s->M, s->x.as<double>(), s->topo.tri.as<uint32_t>(), s->phi.as<double>(), psi, P, dtg,
                        dout);
        d2h(out, dout, sizeof(double) * P, ctx->stream);
    });
}

int bemhip_sim_evolve_rk4(bemhip_sim* s, double dp, int fixdt, double* dt_out) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        bemhip_ctx* ctx = s->ctx;
        const int64_t N = s->N;
        const size_t n3 = 3 * (size_t)N;
        if (s->P.wall_guards && any_negative_x(ctx, N, s->x.as<double>()))   // LinLinSim.cpp:198
            throw Error(BEMHIP_ERR_NEG_X_IN_MIRROR, "negative values");
        double* x = s->x.as<double>();
        double* phi = s->phi.as<double>();
        double* x1 = dev<double>(s->x1, n3);
        double* p1 = dev<double>(s->p1, N);
        double* xs = dev<double>(s->xs, n3);
        double* ps = dev<double>(s->ps, N);
        double* kx = dev<double>(s->kx, 4 * n3);
        double* kp = dev<double>(s->kp, 4 * (size_t)N);
        BH_HIP(hipMemcpyAsync(x1, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
        BH_HIP(hipMemcpyAsync(p1, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, ctx->stream));
        const double t0 = s->time;
        bool conv = true;
        stage_step(s, x1, p1, t0, kx, kp);
        conv = conv && s->last.converged;
        const double dt = fixdt ? dp : get_dt(s, dp, kx, kp);
        const bool frozen = s->P.frozen_stage_time != 0;
        const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
        const double ts[3] = {frozen ? t0 : t0 + 0.5 * dt, frozen ? t0 : t0 + 0.5 * dt, frozen ? t0 : t0 + dt};
        for (int k = 1; k < 4; ++k) {
            axpy_state(ctx, N, cs[k - 1], kx + (k - 1) * n3, x1, xs, kp + (size_t)(k - 1) * N, p1, ps);
            stage_step(s, xs, ps, ts[k - 1], kx + k * n3, kp + (size_t)k * N);
            conv = conv && s->last.converged;
        }
        double* avg = dev<double>(s->avg, n3);
        // pf = p1 + dt/6 (...) (LinLinSim.cpp:238) with p1 as the first stage left
        // it: a pinned stage rewrote its pinned entries (ColocSimPin.cpp:186)
        pin_copy_tail(ctx, N, s->n_pin, p1, phi);
        rk4_final(ctx, N, dt, kx, kp, x1, x, phi, avg, s->P.wall_guards, s->P.wall_eps);
        double* vn = dev<double>(s->vn, n3);
        vertex_normals(ctx, s->topo, x, vn);
        psi_project(ctx, N, avg, vn, s->psi.as<double>());
        BH_HIP(hipStreamSynchronize(ctx->stream));
        s->time = t0 + dt;
        if (dt_out) *dt_out = dt;
        BH_CHECK(conv, BEMHIP_ERR_NOT_CONVERGED, "a stage solve stopped above tolerance");
        if (s->P.wall_guards && any_negative_x(ctx, N, x))   // LinLinSim.cpp:256
            throw Error(BEMHIP_ERR_NEG_X_IN_MIRROR, "negative values");
    });
}

int bemhip_sim_evolve_euler(bemhip_sim* s, double dp, int fixdt, double* dt_out) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        bemhip_ctx* ctx = s->ctx;
        const int64_t N = s->N;
        double* kx = dev<double>(s->kx, 12 * (size_t)N);
        double* kp = dev<double>(s->kp, 4 * (size_t)N);
        stage_step(s, s->x.as<double>(), s->phi.as<double>(), s->time, kx, kp);
        const double dt = fixdt ? dp : get_dt(s, dp, kx, kp);
        euler_update(ctx, N, dt, kx, kp, s->x.as<double>(), s->phi.as<double>());
        BH_HIP(hipStreamSynchronize(ctx->stream));
        s->time += dt;
        if (dt_out) *dt_out = dt;
        BH_CHECK(s->last.converged, BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance");
    });
}

int bemhip_sim_tune(bemhip_sim* s, int32_t cap, bemhip_tune_entry* entries, int32_t* count) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(cap >= 0 && (cap == 0 || entries), BEMHIP_ERR_ARG, "bad tuning output");
        bemhip_ctx* ctx = s->ctx;
        const int solver = s->P.solver;
        const bool batch = s->topo.probs.size() > 1 && solver == BEMHIP_SOLVE_LU16_IR64 && ctx->batch_lu;
        std::vector<bemhip_tune_entry> cand;
        auto add = [&](int nb, int split, int la, int reserve) {
            cand.push_back(bemhip_tune_entry{nb, split, la, reserve, 0.0, 0, 0});
        };
        if (batch) {
            for (int nb : {512, 1024})
                for (int split : {1, 2, 3})
                    for (int la : {1, 0}) add(nb, split, la, ctx->mm_reserve);
        } else if (solver == BEMHIP_SOLVE_LU16_IR64) {
            for (int nb : {768, 1024})
                for (int r : {16, 32, 48}) add(nb, 0, -1, r);
        } else if (solver == BEMHIP_SOLVE_LU16X3_IR64) {
            for (int nb : {768, 1024, 1536}) add(nb, 0, -1, ctx->mm_reserve);
        } else if (solver == BEMHIP_SOLVE_LU32_IR64) {
            for (int nb : {256, 512, 768}) add(nb, 0, -1, ctx->mm_reserve);
        } else if (solver == BEMHIP_SOLVE_LU64) {
            for (int nb : {512, 768, 1024}) add(nb, 0, -1, ctx->mm_reserve);
        }
        // the knobs of a candidate on the context and its solver lanes
        auto apply = [&](const bemhip_tune_entry& c) {
            auto set = [&](Solver* sv) {
                if (solver == BEMHIP_SOLVE_LU16_IR64) sv->nb16 = c.nb;
                else if (solver == BEMHIP_SOLVE_LU16X3_IR64) sv->nb_split = c.nb;
                else if (solver == BEMHIP_SOLVE_LU32_IR64) sv->nb32 = c.nb;
                else if (solver == BEMHIP_SOLVE_LU64) sv->nb64 = c.nb;
                sv->mm_reserve = c.reserve;
            };
            set(ctx);
            for (auto& l : ctx->lanes) set(l.get());
            if (batch) {
                ctx->batch_split = c.split;
                ctx->batch_lookahead = c.lookahead != 0;
            }
        };
        double* hphi = dev<double>(s->hphi, s->N);
        double* psi = s->psi.as<double>();
        hipEvent_t e0 = ctx->ev[2], e1 = ctx->ev[3];
        auto run = [&]() {
            bemhip_solve_stats st{};
            BH_HIP(hipEventRecord(e0, ctx->stream));
            sim_psi(s, s->x.as<double>(), s->phi.as<double>(), hphi, psi, &st);
            BH_HIP(hipEventRecord(e1, ctx->stream));
            BH_CHECK(st.converged, BEMHIP_ERR_NOT_CONVERGED, "solver stopped above tolerance while tuning");
            return (double)event_ms(e0, e1);
        };
        int best = -1;
        if (!cand.empty()) {
            apply(cand[0]);
            (void)run();   // first use: workspaces sized, lanes created
            for (size_t i = 0; i < cand.size(); ++i) {
                apply(cand[i]);
                const double a = run(), b = run();
                cand[i].ms = std::min(a, b);
                if (best < 0 || cand[i].ms < cand[best].ms) best = (int)i;
            }
            cand[best].chosen = 1;
            apply(cand[best]);
        }
        for (int i = 0; i < std::min<int>(cap, (int)cand.size()); ++i) entries[i] = cand[i];
        if (count) *count = (int32_t)cand.size();
    });
}

int bemhip_sim_last_stats(bemhip_sim* s, bemhip_solve_stats* st) {
    if (!s || !st) return BEMHIP_ERR_ARG;
    *st = s->last;
    return BEMHIP_OK;
}

int bemhip_sim_accum_stats(bemhip_sim* s, bemhip_solve_stats* st, int reset) {
    if (!s) return BEMHIP_ERR_ARG;
    if (st) *st = s->accum;
    if (reset) s->accum = bemhip_solve_stats{};
    return BEMHIP_OK;
}

}  // extern "C"

// ----------------------------------------------------------- adjoint --
namespace {

// one RK4 step without wall guards or the psi projection (LinLinSim.cpp:194-241):
// (x, phi) at t0 -> (x, phi) at t0 + dt; the stage slopes stay in s->kx, s->kp
// and the start state in s->x1, s->p1
void rk4_plain(bemhip_sim* s, double* x, double* phi, double t0, double dt) {
    bemhip_ctx* ctx = s->ctx;
    const int64_t N = s->N;
    const size_t n3 = 3 * (size_t)N;
    double* x1 = dev<double>(s->x1, n3);
    double* p1 = dev<double>(s->p1, N);
    double* xs = dev<double>(s->xs, n3);
    double* ps = dev<double>(s->ps, N);
    double* kx = dev<double>(s->kx, 4 * n3);
    double* kp = dev<double>(s->kp, 4 * (size_t)N);
    BH_HIP(hipMemcpyAsync(x1, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(p1, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, ctx->stream));
    const bool frozen = s->P.frozen_stage_time != 0;
    const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
    const double ts[4] = {t0, frozen ? t0 : t0 + 0.5 * dt, frozen ? t0 : t0 + 0.5 * dt, frozen ? t0 : t0 + dt};
    stage(s, x1, p1, ts[0], kx, kp);
    BH_CHECK(s->last.converged, BEMHIP_ERR_NOT_CONVERGED, "a stage solve stopped above tolerance");
    for (int k = 1; k < 4; ++k) {
        axpy_state(ctx, N, cs[k - 1], kx + (k - 1) * n3, x1, xs, kp + (size_t)(k - 1) * N, p1, ps);
        stage(s, xs, ps, ts[k], kx + k * n3, kp + (size_t)k * N);
        BH_CHECK(s->last.converged, BEMHIP_ERR_NOT_CONVERGED, "a stage solve stopped above tolerance");
    }
    double* avg = dev<double>(s->avg, n3);
    rk4_final(ctx, N, dt, kx, kp, x1, x, phi, avg, 0, 0.0);
}

double total_volume(bemhip_sim* s, const double* x) {
    double v = 0.0;
    for (double a : part_volumes_host(s, x)) v += a;
    return v;
}

}  // namespace

extern "C" {

int bemhip_sim_adjoint_radius(bemhip_sim* s, int32_t nsteps, double dt, const double* target, int32_t every,
                              double* loss, double* radius, double* grad_x, double* grad_phi) {
    if (!s) return BEMHIP_ERR_ARG;
    return guarded(s->ctx, [&] {
        BH_CHECK(target && loss && nsteps > 0 && dt > 0.0 && every > 0, BEMHIP_ERR_ARG, "bad arguments");
        BH_CHECK(s->P.sigma == 0.0, BEMHIP_ERR_ARG, "the adjoint supports sigma = 0 only");
        BH_CHECK(!(s->P.flags & (BEMHIP_MIRROR_X0 | BEMHIP_INDEPENDENT_PARTS)) && !s->P.wall_guards, BEMHIP_ERR_ARG,
                 "the adjoint supports the free-space, coupled problem only");
        BH_CHECK(s->P.solver != BEMHIP_SOLVE_BICGSTAB64, BEMHIP_ERR_ARG, "the adjoint needs an LU-based solver");
        BH_CHECK(s->n_pin == 0, BEMHIP_ERR_ARG, "the adjoint does not support pinned vertices");
        bemhip_ctx* ctx = s->ctx;
        hipStream_t st = ctx->stream;
        const int64_t N = s->N;
        const size_t n3 = 3 * (size_t)N;
        const double t0 = s->time;
        double* x = s->x.as<double>();
        double* phi = s->phi.as<double>();
        double* x0 = dev<double>(s->a_x0, n3);
        double* p0 = dev<double>(s->a_p0, N);
        BH_HIP(hipMemcpyAsync(x0, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
        BH_HIP(hipMemcpyAsync(p0, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        const int nck = (nsteps + every - 1) / every;
        double* ckx = dev<double>(s->a_ckx, n3 * nck);
        double* ckp = dev<double>(s->a_ckp, (size_t)N * nck);
        std::vector<double> R(nsteps), V(nsteps);

        // forward sweep with a checkpoint every `every` steps
        double L = 0.0;
        for (int n = 0; n < nsteps; ++n) {
            if (n % every == 0) {
                BH_HIP(hipMemcpyAsync(ckx + n3 * (n / every), x, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                BH_HIP(hipMemcpyAsync(ckp + (size_t)N * (n / every), phi, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                      st));
            }
            rk4_plain(s, x, phi, t0 + n * dt, dt);
            V[n] = total_volume(s, x);
            R[n] = std::cbrt(3.0 * V[n] / (4.0 * PI));   // volume-equivalent radius
            L += (R[n] - target[n]) * (R[n] - target[n]);
        }

        // reverse sweep, segment by segment
        double* xb = dev<double>(s->a_xb, n3);
        double* pb = dev<double>(s->a_pb, N);
        BH_HIP(hipMemsetAsync(xb, 0, sizeof(double) * n3, st));
        BH_HIP(hipMemsetAsync(pb, 0, sizeof(double) * N, st));
        double* sgx = dev<double>(s->a_sgx, n3 * every);
        double* sgp = dev<double>(s->a_sgp, (size_t)N * every);
        double* kxb = dev<double>(s->a_kxb, 4 * n3);
        double* kpb = dev<double>(s->a_kpb, 4 * (size_t)N);
        double* xsb = dev<double>(s->a_xsb, n3);
        double* psb = dev<double>(s->a_psb, N);
        double* psib = dev<double>(s->a_psib, N);
        double* lam = dev<double>(s->a_lam, N);
        double* kxt = dev<double>(s->a_kxt, n3);
        double* kpt = dev<double>(s->a_kpt, N);
        const int np = s->topo.nparts;
        double* vc = dev<double>(s->a_vc, np);
        const bool frozen = s->P.frozen_stage_time != 0;
        const bool mixed_lu = mixed(s->P.solver);
        StageParams sp = stage_params(s);
        for (int c = nck - 1; c >= 0; --c) {
            const int n0 = c * every, n1 = std::min(nsteps, n0 + every);
            BH_HIP(hipMemcpyAsync(x, ckx + n3 * c, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
            BH_HIP(hipMemcpyAsync(phi, ckp + (size_t)N * c, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            for (int n = n0; n < n1; ++n) {
                BH_HIP(hipMemcpyAsync(sgx + n3 * (n - n0), x, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                BH_HIP(hipMemcpyAsync(sgp + (size_t)N * (n - n0), phi, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                      st));
                if (n + 1 < n1) rk4_plain(s, x, phi, t0 + n * dt, dt);
            }
            for (int n = n1 - 1; n >= n0; --n) {
                // step n again from its start state: slopes in s->kx / s->kp, x_{n+1} in x
                BH_HIP(hipMemcpyAsync(x, sgx + n3 * (n - n0), sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                BH_HIP(hipMemcpyAsync(phi, sgp + (size_t)N * (n - n0), sizeof(double) * N,
                                      hipMemcpyDeviceToDevice, st));
                const double tn = t0 + n * dt;
                rk4_plain(s, x, phi, tn, dt);
                // loss term of x_{n+1}: d(R - target)^2/dV = 2 (R - target) R / (3 V)
                const double cf = 2.0 * (R[n] - target[n]) * R[n] / (3.0 * V[n]);
                std::vector<double> h(np, cf);
                h2d(vc, h.data(), sizeof(double) * np, st);
                adj_volume(ctx, s->topo, x, vc, xb);
                // slope cotangents (dt/6) (1, 2, 2, 1) of the RK4 combination
                const double wgt[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
                for (int k = 0; k < 4; ++k) axpy_pair(ctx, N, wgt[k], xb, kxb + k * n3, pb, kpb + (size_t)k * N, false);
                const double* x1 = s->x1.as<double>();
                const double* p1 = s->p1.as<double>();
                const double* kx = s->kx.as<double>();
                const double* kp = s->kp.as<double>();
                double* xs = s->xs.as<double>();
                double* ps = s->ps.as<double>();
                const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
                const double ts[4] = {tn, frozen ? tn : tn + 0.5 * dt, frozen ? tn : tn + 0.5 * dt,
                                      frozen ? tn : tn + dt};
                for (int k = 3; k >= 0; --k) {
                    const double* xk = x1;
                    const double* pk = p1;
                    if (k > 0) {
                        axpy_state(ctx, N, cs[k - 1], kx + (k - 1) * n3, x1, xs, kp + (size_t)(k - 1) * N, p1, ps);
                        xk = xs;
                        pk = ps;
                    }
                    // the stage again: its LU factors, psi and part volumes
                    stage(s, xk, pk, ts[k], kxt, kpt);
                    AdjStageIn in{xk, pk, s->psil.as<double>(), kxt, s->pvol.as<double>(), s->pphys.as<double>(),
                                  sp.epsilon, sp.gamma};
                    adj_stage_pre(ctx, s->topo, in, kxb + k * n3, kpb + (size_t)k * N, psib);
                    bemhip_solve_stats tst{};
                    solve_transposed_device(ctx, (int)N, ctx->G64.as<double>(),
                                            mixed_lu ? ctx->G32.as<float>() : nullptr, s->P.solver, psib, lam,
                                            s->P.tol, s->P.maxit, &tst);
                    BH_CHECK(tst.converged, BEMHIP_ERR_NOT_CONVERGED, "transposed solve stopped above tolerance");
                    BH_HIP(hipMemsetAsync(xsb, 0, sizeof(double) * n3, st));
                    BH_HIP(hipMemsetAsync(psb, 0, sizeof(double) * N, st));
                    adj_stage_post(ctx, s->topo, in, lam, xsb, psb);
                    // the stage input is x1 + c kx_{k-1}: both receive its cotangent
                    axpy_pair(ctx, N, 1.0, xsb, xb, psb, pb, true);
                    if (k > 0) axpy_pair(ctx, N, cs[k - 1], xsb, kxb + (k - 1) * n3, psb, kpb + (size_t)(k - 1) * N, true);
                }
            }
        }
        if (grad_x) d2h(grad_x, xb, sizeof(double) * n3, st);
        if (grad_phi) d2h(grad_phi, pb, sizeof(double) * N, st);
        if (radius) std::copy(R.begin(), R.end(), radius);
        *loss = L;
        // leave the simulation where it started
        BH_HIP(hipMemcpyAsync(x, x0, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
        BH_HIP(hipMemcpyAsync(phi, p0, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        BH_HIP(hipStreamSynchronize(st));
        s->time = t0;
    });
}

}  // extern "C"
