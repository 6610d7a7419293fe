/*
 * bem_oracle.h -- FP64 CPU restatement of the reference's per-RK4-stage BEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and the benchmark's CPU-baseline leg may load it; the product
 * path (libbemhip) never links or calls it.
 *
 * Conventions (shared with the reference, Bem/basic/Bem.hpp:10-16):
 *   - xyz   : N x 3 doubles, array of structs (vec3 = {x, y, z})
 *   - tri   : M x 3 vertex indices, right-handed (outward) order
 *   - G, H  : N x N doubles, column-major (Eigen::MatrixXd), G(i,j) = G[j*N+i]
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef BEM_ORACLE_H
#define BEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quadrature tables used by the collocation path (quadrature.hpp:30-38, 88-96) */
void bo_quadrature(double* gauss7_x, double* gauss7_w,      /* 7 + 7 on (0,1)        */
                   double* tri7_x, double* tri7_y, double* tri7_w); /* 7 x 3 on unit triangle */

/* the 19-point degree-9 rule of the exterior potential (quadrature.hpp:126-146) */
void bo_quadrature19(double* tri19_x, double* tri19_y, double* tri19_w);
/* one disjoint panel under the 19-point rule (Integrator.hpp:471-491) */
void bo_exterior_panel(const double* y, const double* a, const double* b, const double* c, double* g3, double* h3);
/* compute_exterior_pot (LinLinSim.cpp:440-477): out[P] at the points pts (P x 3) */
void bo_exterior_pot(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, const double* phi,
                     const double* psi, int64_t P, const double* pts, double* out, int nthreads);

/* single (collocation point, triangle) panel integrals, 3 G + 3 H values in
 * triangle-vertex order (Integrator.cpp:237-283) */
void bo_panel(const double* x, const double* a, const double* b, const double* c,
              int self_vertex /* -1 disjoint, else 0/1/2 = x is that vertex */,
              int mirror, double* g3, double* h3);

/* ColocSim::assemble_matrices (ColocSim.cpp:21-110).  G, H may not be NULL. */
void bo_assemble(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, int mirror,
                 double* G, double* H, int nthreads);

/* dense helpers (column-major) */
void bo_gemv(int64_t N, const double* A, const double* x, double* y, int nthreads);
/* Eigen::PartialPivLU compute+solve restated as a blocked right-looking GEPP
 * (Simulation.cpp:57-61).  A is overwritten by its factors; piv receives the
 * row interchanges (LAPACK style).  Returns 0 or the 1-based index of the
 * first exactly-zero pivot. */
int64_t bo_lu_factor(int64_t N, double* A, int64_t* piv, int nthreads);
void bo_lu_solve(int64_t N, const double* LU, const int64_t* piv, double* b);
/* Eigen::BiCGSTAB<MatrixXd> with its default diagonal preconditioner
 * (Simulation.cpp:48-56).  tol <= 0 means Eigen's default (machine epsilon),
 * maxit <= 0 means 2N.  Returns the iteration count, *err = |r|/|b|. */
int64_t bo_bicgstab(int64_t N, const double* A, const double* b, double* x,
                    double tol, int64_t maxit, double* err, int nthreads);

/* mesh geometry (Mesh.cpp) */
void bo_triangle_normals(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* n);
void bo_vertex_normals(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* n);
void bo_tangent_gradients(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                          const double* pot, double* grad);
void bo_curvatures(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                   double* kappa_tri, double* gauss_tri);
void bo_kappa(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* kappa_vert);
/* connected components (split_by_loose_parts, Mesh.cpp:382-440): comp[v] in
 * order of the smallest vertex of each part; returns the part count */
int64_t bo_components(int64_t N, int64_t M, const uint32_t* tri, int64_t* comp);
/* signed volume per part (Mesh.cpp:206-218), vol has room for the part count */
void bo_part_volumes(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                     const int64_t* comp, int64_t nparts, double* vol);

/* simulation parameters (Simulation.hpp:29-45, LinLinSim.hpp:22-27) */
typedef struct {
    double p_inf, epsilon, sigma, gamma;
    double V_0;
    double field_amp, field_freq, field_phase;  /* p_field = amp sin(2 pi f t + phase) */
    double dp_balance, min_dt;
    double wall_eps;                             /* nopenetration eps (LinLinSim.hpp:24) */
    int mirror;                                  /* MIRROR_MESH */
    int wall_guards;                             /* test_negative + nopenetration */
    int frozen_stage_time;                       /* 1 = reference (all stages at step start) */
    int solver;                                  /* 0 = PartialPivLU, 1 = BiCGSTAB */
    int nthreads;
    int64_t n_pin;                               /* ColocSimPin: the last n_pin vertices are pinned */
} bo_params;

/* ColocSimPin::set_x_boundary (ColocSimPin.cpp:222-287): vertices with
 * |x| < 1e-5 get x = 0 and move to the end (swap with the last unvisited
 * vertex), triangles are relabelled; phi / psi (may be NULL) move with their
 * vertices.  Returns the number of pinned vertices. */
int64_t bo_pin_boundary(int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi, double* psi);
/* the contact-line drift law (ColocSimPin.cpp:126-136) */
double bo_drift_velocity(double costheta);

/* LinLinSim::position_t (LinLinSim.cpp:26-83): vertex velocities (N x 3); psi
 * (N) receives the solved normal derivative. */
void bo_position_t(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                   const uint32_t* tri, const double* pot, double* vel, double* psi);
/* LinLinSim::pot_t_multi (LinLinSim.cpp:98-120) */
void bo_pot_t_multi(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                    const uint32_t* tri, const double* vel, double t, double* dpot);
/* LinLinSim::evolve_system_RK4 (LinLinSim.cpp:194-257).  xyz, phi, psi, time
 * are updated in place; returns 0, or -1 when a wall-mode vertex has x < 0
 * (the reference throws std::out_of_range). *dt_out receives dt. */
/* cache-blocked GEPP (PartialPivLU's structure: GEPP panels of NB columns, packed GEMM
 * trailing updates); max_panels > 0 stops after that many panels (a timed sample) */
int64_t bo_lu_factor_blocked(int64_t N, double* A, int64_t* piv, int nthreads, int64_t NB, int64_t max_panels);
/* Simulation::get_dt (Simulation.cpp:77-92): dp / (max|dpot| + dp_balance max|vel|), capped by min_dt > 0 */
double bo_get_dt(const bo_params* P, int64_t N, double dp, const double* vel, const double* dpot);
/* LinLinSim::nopenetration (LinLinSim.cpp:162-190): xf from x0 and the velocities c (N x 3); rewrites c.x
 * of the vertices it slows down */
void bo_nopenetration(int64_t N, double eps, double dt, const double* x0, double* c, double* xf);
int bo_rk4_step(const bo_params* P, int64_t N, double* xyz, int64_t M, const uint32_t* tri,
                double* phi, double* psi, double* time, double dp, int fixdt, double* dt_out);
/* LinLinSim::evolve_system (Euler, LinLinSim.cpp:260-284); psi (the state's
 * psi, read at pinned vertices only) may be NULL when n_pin = 0 */
int bo_euler_step(const bo_params* P, int64_t N, double* xyz, int64_t M, const uint32_t* tri,
                  double* phi, const double* psi, double* time, double dp, int fixdt, double* dt_out);
/* Simulation::compute_psi (Simulation.hpp:73-77) */
void bo_compute_psi(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                    const uint32_t* tri, const double* phi, double* psi);

#ifdef __cplusplus
}
#endif
#endif
