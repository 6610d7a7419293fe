/*
 * bemhip.h -- C-ABI of the MI355X (gfx950) per-RK4-stage BEM path.
 *
 * Plain C types only (pointers, sizes, POD structs); no exceptions cross this
 * boundary.  Every entry point returns a bemhip_status; the message of the last
 * failure on a context is available from bemhip_last_error().
 *
 * Layout conventions (identical to the reference's, Bem/basic/Bem.hpp:10-16):
 *   xyz  N x 3 doubles, array of structs {x, y, z} (std::vector<vec3>)
 *   tri  M x 3 uint32 vertex indices, right-handed (outward) order
 *        (std::vector<Triplet> holds size_t; the host C++ class narrows them)
 *   G, H N x N doubles, column-major (Eigen::MatrixXd), G(i,j) = G[j*N + i]
 *
 * Host buffers passed in are caller-owned and copied; device state is owned by
 * the context / simulation handle.  One host thread per context; calls are
 * issued on the context's stream and synchronous at return.
 */
#ifndef BEMHIP_H
#define BEMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BEMHIP_OK = 0,
    BEMHIP_ERR_ARG = 1,             /* invalid argument (sizes, null pointers, flags)  */
    BEMHIP_ERR_BAD_MESH = 2,        /* index out of range, degenerate triangle          */
    BEMHIP_ERR_NEG_X_IN_MIRROR = 3, /* wall mode with a vertex at x < 0 (LinLinSim.cpp:148-159 throws) */
    BEMHIP_ERR_SINGULAR = 4,        /* exact zero pivot in the LU                       */
    BEMHIP_ERR_NOT_CONVERGED = 5,   /* refinement / BiCGSTAB stopped above tolerance    */
    BEMHIP_ERR_NAN = 6,             /* non-finite value in inputs or results            */
    BEMHIP_ERR_OOM = 7,             /* device allocation failed                         */
    BEMHIP_ERR_HIP = 8,             /* HIP / rocBLAS runtime error                      */
    BEMHIP_ERR_NO_DEVICE = 9,       /* no usable gfx950 device                          */
    BEMHIP_ERR_IO = 10              /* file missing, unreadable or not a supported PLY  */
} bemhip_status;

/* assembly flags */
#define BEMHIP_MIRROR_X0 0x1u   /* image across the wall x = 0 (MIRROR_MESH, LinLinSim.hpp:13) */
/* Batch of independent bubbles: every loose part of the mesh is its own BEM
 * problem (its own dense G, solve, V_0 and forcing), as if each part were a
 * separate reference Simulation advanced with a shared dt.  Parts must be
 * contiguous vertex ranges in increasing order.  Accepted by
 * bemhip_compute_psi and in bemhip_sim_params.flags; not by bemhip_assemble
 * (whose output is one N x N matrix). */
#define BEMHIP_INDEPENDENT_PARTS 0x2u

/* dense solvers for G psi = H phi */
typedef enum {
    BEMHIP_SOLVE_LU64 = 0,        /* FP64 blocked GEPP (Eigen::PartialPivLU semantics)          */
    BEMHIP_SOLVE_LU32_IR64 = 1,   /* FP32 GEPP factor + FP64-residual iterative refinement       */
    BEMHIP_SOLVE_BICGSTAB64 = 2,  /* FP64 BiCGSTAB, diagonal preconditioner (Eigen::BiCGSTAB)    */
    BEMHIP_SOLVE_LU16X3_IR64 = 3, /* FP32 GEPP whose trailing updates run as split-fp16 x3 MFMA
                                     products with FP32 accumulation, + FP64 refinement          */
    BEMHIP_SOLVE_LU16_IR64 = 4    /* GEPP with every block product (trailing updates, L11^-1 A12,
                                     A21 U11^-1) on fp16 MFMA with FP32 accumulation and an FP32
                                     factor, + FP64 refinement (2-3 sweeps); falls back to FP32
                                     products if an operand leaves the fp16 range                */
} bemhip_solver;
/* Refinement of the mixed solvers: x = LU \ b, then sweeps
 * x += LU \ (b - G x) with the FP64 residual.  The loop stops before a sweep
 * once |b - G x|_inf <= 1e-10 |b|_inf, or after a sweep once the correction
 * |d|_inf <= tol |x|_inf (tol <= 0: 1e-9); `iterations` counts the sweeps. */

typedef struct bemhip_ctx bemhip_ctx;
typedef struct bemhip_sim bemhip_sim;

typedef struct {
    int32_t iterations;      /* refinement sweeps or BiCGSTAB iterations           */
    int32_t converged;       /* 1 when the stopping test was met                   */
    double residual;         /* |G psi - H phi|_2 / |H phi|_2 in FP64              */
    double correction;       /* last refinement sweep's |d|_inf / |psi|_inf        */
    double ms_assemble;      /* device time of the assembly kernels                */
    double ms_factor;        /* device time of the factorisation                   */
    double ms_solve;         /* device time of solves / refinement / iterations    */
    double gemm_ms;          /* device time inside trailing-update GEMMs           */
    double gemm_flops;       /* algorithmic flops of those GEMMs (2 m n k each)    */
    int32_t gemm_launches;   /* number of trailing-update GEMM launches            */
    int32_t pad_;
    int32_t panels;          /* LU panels factored                                 */
    int32_t panels_pivoted;  /* panels that needed row interchanges                */
} bemhip_solve_stats;

/* physics and flags of a simulation (Simulation.hpp:29-45, LinLinSim.hpp:22-27) */
typedef struct {
    double p_inf, epsilon, sigma, gamma;
    double V_0;                 /* NaN: total initial volume (Simulation.hpp:42) */
    double field_amp, field_freq, field_phase;  /* p_field = amp sin(2 pi freq t + phase) */
    double dp_balance;          /* get_dt balance (Simulation.hpp:37), default 3 */
    double min_dt;              /* <= 0: unused (Simulation.cpp:86) */
    double wall_eps;            /* nopenetration eps (LinLinSim.hpp:24), default 1e-2 */
    double tol;                 /* solver tolerance (<= 0: method default) */
    int32_t maxit;              /* solver iteration cap (<= 0: method default) */
    int32_t solver;             /* bemhip_solver; the reference default is BiCGSTAB */
    uint32_t flags;             /* BEMHIP_MIRROR_X0 | BEMHIP_INDEPENDENT_PARTS */
    int32_t wall_guards;        /* test_negative + nopenetration (reference RK4 in wall mode) */
    int32_t frozen_stage_time;  /* 1: every stage uses the step-start time (LinLinSim.cpp:208-238) */
    int32_t pad_;
} bemhip_sim_params;

/* ---- context ------------------------------------------------------------ */
int bemhip_create(int device, bemhip_ctx** out);
int bemhip_destroy(bemhip_ctx* ctx);
const char* bemhip_last_error(const bemhip_ctx* ctx);
/* quadrature constants the kernels use: gauss7 x,w (7+7, on (0,1)) then the
 * triangle rule x,y,w (7+7+7); 35 doubles (quadrature.hpp:30-38, 88-96).
 * Needs no device. */
void bemhip_quadrature(double* out35);

/* ---- ColocSim::assemble_matrices (ColocSim.cpp:21-110) ------------------
 * G and H, when non-null, receive the column-major N x N matrices of the
 * reference, H including its 4-pi diagonal.  When phi and Hphi are both
 * non-null, Hphi = H phi is formed inside the assembly without storing H
 * (then G and H may both be null). */
int bemhip_assemble(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                    uint32_t flags, double* G, double* H, const double* phi, double* Hphi);

/* ---- Simulation::solve_system (Simulation.cpp:40-71) --------------------
 * Solves G x = b for a column-major FP64 G given on the host. */
int bemhip_solve(bemhip_ctx* ctx, int64_t N, const double* G, const double* b, double* x,
                 int solver, double tol, int32_t maxit, bemhip_solve_stats* stats);

/* ---- Simulation::compute_psi (Simulation.hpp:73-77) ---------------------
 * Assembly + H phi + solve, device resident; psi receives dphi/dn. */
int bemhip_compute_psi(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                       const double* phi, uint32_t flags, int solver, double tol, int32_t maxit,
                       double* psi, bemhip_solve_stats* stats);

/* ---- simulation handle (LinLinSim / ColocSim state, device resident) ---- */
void bemhip_sim_default_params(bemhip_sim_params* p);
int bemhip_sim_create(bemhip_ctx* ctx, const bemhip_sim_params* p, int64_t N, const double* xyz,
                      int64_t M, const uint32_t* tri, bemhip_sim** out);
int bemhip_sim_destroy(bemhip_sim* sim);   /* before bemhip_destroy of its context */
int bemhip_sim_set_phi(bemhip_sim* sim, const double* phi);            /* Simulation.hpp:92-95   */
int bemhip_sim_set_psi(bemhip_sim* sim, const double* psi);            /* Simulation.hpp:97-100  */
int bemhip_sim_set_V_0(bemhip_sim* sim, double V_0);                   /* Simulation.hpp:106-108 */
int bemhip_sim_set_time(bemhip_sim* sim, double t);                    /* Simulation.hpp:120-122 */
int bemhip_sim_set_solver(bemhip_sim* sim, int solver, double tol, int32_t maxit);
/* replace the physics / solver / flag parameters (set_min_dt, set_dp_balance,
 * set_bcgstab, set_V_0 of Simulation.hpp:106-159); a NaN V_0 keeps the current one */
int bemhip_sim_set_params(bemhip_sim* sim, const bemhip_sim_params* p);
/* new mesh after remeshing (LinLinSim::remesh, LinLinSim.cpp:385-438): the
 * topology tables are rebuilt, phi and psi are reset to zero and must be set
 * again; time and V_0 are kept */
int bemhip_sim_set_mesh(bemhip_sim* sim, int64_t N, const double* xyz, int64_t M, const uint32_t* tri);
int bemhip_sim_get_state(bemhip_sim* sim, double* xyz, double* phi, double* psi, double* time);
int bemhip_sim_get_volume(bemhip_sim* sim, double* vol);               /* Simulation.hpp:128-130 */
/* the triangles in the simulation's vertex numbering (the public mesh.trigs) */
int bemhip_sim_get_triangles(bemhip_sim* sim, uint32_t* tri);
/* Simulation::export_mesh (Simulation.cpp:94-96; float, precision 4) and its
 * lossless double variant (precision 8): see bemhip_ply_write */
int bemhip_sim_export_ply(bemhip_sim* sim, const char* path, int32_t precision);
/* number of loose parts (split_by_loose_parts, Mesh.cpp:382-440) and their
 * signed volumes, part p = the part holding the smallest vertex id not in
 * parts 0..p-1 */
int bemhip_sim_get_parts(bemhip_sim* sim, int32_t* nparts);
int bemhip_sim_get_part_volumes(bemhip_sim* sim, double* vol, int32_t nparts);
/* Per-part Bernoulli parameters for pot_t_multi (LinLinSim.cpp:98-120): part p
 * uses V_0[p] and p_field = field_amp[p] sin(2 pi field_freq[p] t +
 * field_phase[p]).  A null array leaves that parameter as it is.  No
 * counterpart in the reference, whose V_0 and forcing are one scalar per
 * Simulation; it serves batches of independent bubbles
 * (BEMHIP_INDEPENDENT_PARTS), where each part's V_0 defaults to its own initial
 * volume.  bemhip_sim_set_V_0, and bemhip_sim_set_params with a non-NaN V_0,
 * reset every part to the scalar V_0; set_params resets nothing else. */
int bemhip_sim_set_part_physics(bemhip_sim* sim, int32_t nparts, const double* V_0, const double* field_amp,
                                const double* field_freq, const double* field_phase);
/* LinLinSim::position_t (LinLinSim.cpp:26-83) on the current state: vel N x 3
 * and psi N (either may be null) */
int bemhip_sim_position_t(bemhip_sim* sim, double* vel, double* psi);
/* LinLinSim::pot_t_multi (LinLinSim.cpp:98-120) for given velocities */
int bemhip_sim_pot_t_multi(bemhip_sim* sim, const double* vel, double t, double* dpot);
/* Simulation::compute_psi on the current state */
int bemhip_sim_compute_psi(bemhip_sim* sim);
/* LinLinSim::evolve_system_RK4 (LinLinSim.cpp:194-257); *dt_out may be null */
int bemhip_sim_evolve_rk4(bemhip_sim* sim, double dp, int fixdt, double* dt_out);
/* LinLinSim::evolve_system, Euler (LinLinSim.cpp:260-284) */
int bemhip_sim_evolve_euler(bemhip_sim* sim, double dp, int fixdt, double* dt_out);
/* statistics of the last stage's solve */
int bemhip_sim_last_stats(bemhip_sim* sim, bemhip_solve_stats* stats);
/* sums of the timing fields (and iteration counts) over all stages since the
 * last reset; residual / correction hold the maxima */
int bemhip_sim_accum_stats(bemhip_sim* sim, bemhip_solve_stats* stats, int reset);

/* ---- solver tuning (no counterpart in the reference) ---------------------
 * Times the simulation's assembly + solve at its current state (best of two
 * runs each) under every candidate setting of the block size of the solver
 * in use, and for a batch of independent bubbles solved with lu16_ir64 also
 * of its schedule (sub-classes per size class, look-ahead inside a class);
 * for a single fp16 LU also the compute units left to its panel stream.  The
 * fastest setting stays on the context (for all of its simulations), the
 * simulation's state is unchanged (psi receives the last run's solution).
 * entries (up to cap, may be null) receive every candidate; *count the number
 * of candidates. */
typedef struct {
    int32_t nb;          /* LU outer block of the solver in use              */
    int32_t split;       /* batch: sub-classes per size class (0: n/a)       */
    int32_t lookahead;   /* batch: look-ahead inside a class (-1: n/a)       */
    int32_t reserve;     /* fp16 LU: compute units left to the panel stream  */
    double ms;           /* assembly + solve, best of two runs               */
    int32_t chosen;      /* 1 for the setting kept                           */
    int32_t pad_;
} bemhip_tune_entry;
int bemhip_sim_tune(bemhip_sim* sim, int32_t cap, bemhip_tune_entry* entries, int32_t* count);

/* ---- potential off the surface ----
 * phi_ext(y) = 1/(4 pi) sum_T [ -sum_k G_k(y,T) psi_k + sum_k H_k(y,T) phi_k ]
 * for P points y (points: P x 3), with the disjoint-panel integrals under the
 * 19-point degree-9 rule and no mirror image; phi and psi hold N values.
 * Replaces the free function compute_exterior_pot (LinLinSim.cpp:440-477,
 * declared LinLinSim.hpp:17; per triangle Integrator.cpp:287-305).  Points
 * on the surface are singular, as in the reference. */
int bemhip_exterior_potential(bemhip_ctx* ctx, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                              const double* phi, const double* psi, int64_t P, const double* points,
                              double* out);
/* ---- No-allocation variant of bemhip_exterior_potential -----------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the same potential, but every array is a device pointer owned
 * by the caller (d_xyz N x 3, d_tri M x 3, d_phi and d_psi N, d_points P x 3,
 * d_out P), and the work buffers (triangle records, per-chunk partial sums,
 * an error word) live in a fixed 64 MiB device array of the library, one per
 * GPU: 104 M + P <= 8388576 (e.g. M <= 65536 triangles with P up to 1.5
 * million points), BEMHIP_ERR_ARG beyond.  Indices and finiteness are checked
 * on the device (BEMHIP_ERR_BAD_MESH, BEMHIP_ERR_NAN).  Runs on the context's
 * stream, synchronous at return; calls no hipMalloc / hipFree.  Calls that
 * overlap on one GPU (from any context) share the array and race. */
int bemhip_exterior_potential_noalloc(bemhip_ctx* ctx, int64_t N, const double* d_xyz, int64_t M,
                                      const uint32_t* d_tri, const double* d_phi, const double* d_psi, int64_t P,
                                      const double* d_points, double* d_out);
/* LinLinSim::exterior_pot (LinLinSim.cpp:478-483): psi for the current mesh
 * and phi (the simulation's solver and mirror flag; the simulation's psi is
 * left unchanged), then the potential at the P points from (phi, psi). */
int bemhip_sim_exterior_pot(bemhip_sim* sim, int64_t P, const double* points, double* out);

/* ---- pinned contact line (ColocSimPin, ColocSimPin.hpp:13-68) -----------
 * A bubble sitting on the wall x = 0 whose contact-line vertices are pinned to
 * the wall: they are the last n_pin vertices, psi is prescribed there and phi
 * is solved for (the pinned columns of G and H trade places after the 4-pi
 * rule, ColocSimPin.cpp:113-117), and they move along the wall with the drift
 * law of the contact angle (ColocSimPin.cpp:126-136, 176-188).  Wall mode
 * (BEMHIP_MIRROR_X0) is what the reference compiles it with.
 *
 * bemhip_pin_boundary: ColocSimPin::set_x_boundary (ColocSimPin.cpp:222-287)
 * on host arrays, in place: vertices with |x| < 1e-5 get x = 0 and move to the
 * end, triangles are relabelled, phi / psi (may be null) move along; *npin
 * receives their number.  Needs no device. */
int bemhip_pin_boundary(int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi, double* psi,
                        int64_t* npin);
/* the two ColocSimPin constructors (ColocSimPin.hpp:17-52): the mesh (and phi,
 * psi when non-null; zero otherwise) reordered by bemhip_pin_boundary, V_0
 * (when NaN in p) from the mesh as given, before the reordering; *npin (may be
 * null) receives the pinned count.  Not with BEMHIP_INDEPENDENT_PARTS. */
int bemhip_sim_create_pinned(bemhip_ctx* ctx, const bemhip_sim_params* p, int64_t N, const double* xyz,
                             int64_t M, const uint32_t* tri, const double* phi, const double* psi,
                             bemhip_sim** out, int64_t* npin);
/* the public member N_pin (ColocSimPin.hpp:57): set after bemhip_sim_set_mesh
 * with a mesh whose last npin vertices are the pinned ones (0 = ColocSim) */
int bemhip_sim_set_pinned(bemhip_sim* sim, int64_t npin);
int bemhip_sim_get_pinned(bemhip_sim* sim, int64_t* npin);

/* ---- state files: binary little-endian PLY (MeshIO.cpp) ------------------
 * Need no device and no context; on failure bemhip_io_last_error() (per host
 * thread) holds the message.
 * bemhip_ply_write: export_ply_double (precision 8) / export_ply_float
 * (precision 4) with phi and psi (MeshIO.cpp:263-335), with one value named w
 * when psi is null (MeshIO.cpp:191-261), or the bare mesh when both are null
 * (MeshIO.cpp:67-100).  Simulation::export_mesh (Simulation.cpp:94-96) writes
 * precision 4, which rounds the state; precision 8 restores it exactly.
 * bemhip_ply_info / bemhip_ply_read: import_ply (MeshIO.cpp:701-903): counts,
 * the number of per-vertex values (0, 1 = phi, 2 = phi and psi) and their
 * precision, then the data (phi / psi may be null; missing values read as 0). */
const char* bemhip_io_last_error(void);
int bemhip_ply_write(const char* path, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                     const double* phi, const double* psi, int32_t precision);
int bemhip_ply_info(const char* path, int64_t* N, int64_t* M, int32_t* nvalues, int32_t* precision);
int bemhip_ply_read(const char* path, int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi,
                    double* psi);

/* ---- adjoint of a radius trajectory (no counterpart in the reference) ----
 * From the simulation's current state (x0, phi0, time), runs `nsteps` RK4
 * steps of fixed size dt (LinLinSim::evolve_system_RK4 with fixdt, no wall)
 * and evaluates L = sum_{n=1..nsteps} (R_n - target[n-1])^2 with R_n the
 * volume-equivalent radius (3 V_n / 4 pi)^(1/3) of the whole mesh after step
 * n (radius.cpp:45).  The reverse sweep recomputes each segment of `every`
 * steps from a checkpoint, re-assembles and re-factors every stage and solves
 * the transposed systems with those LU factors; no dG/dx is formed.  Outputs
 * L, the radii (nsteps, may be null) and dL/dx0 (N x 3) and dL/dphi0 (N)
 * (either may be null), with V_0, the forcing and the mesh connectivity held
 * fixed.  Needs sigma = 0, no wall mode, the coupled problem and an LU-based
 * solver.  The simulation's state and time are left unchanged. */
int bemhip_sim_adjoint_radius(bemhip_sim* sim, int32_t nsteps, double dt, const double* target, int32_t every,
                              double* loss, double* radius, double* grad_x, double* grad_phi);

#ifdef __cplusplus
}
#endif
#endif
