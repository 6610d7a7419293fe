// Per-stage O(N + M) kernels around the dense solve: triangle geometry,
// vertex normals, curvature, part volumes, vertex velocities, the Bernoulli
// right-hand side and the fused RK4 state updates.  All FP64, one thread per
// triangle or vertex, bandwidth/latency bound.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "quad_rules.hpp"
#include "vec3.hpp"

namespace bemhip {

// generate_triangle_normals (Mesh.cpp:130-138), generate_tangent_gradients
// (LinLinSim.cpp:288-305), signed tetrahedron volume (Mesh.cpp:206-218)
__global__ __launch_bounds__(256) void k_tri_geom(int M, const double* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                  const double* __restrict__ pot, double* __restrict__ tnorm,
                                                  double* __restrict__ tgrad, double* __restrict__ tvol) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t ia = tri[3 * j], ib = tri[3 * j + 1], ic = tri[3 * j + 2];
    const d3 a = ld3(xyz + 3 * (size_t)ia), b = ld3(xyz + 3 * (size_t)ib), c = ld3(xyz + 3 * (size_t)ic);
    if (tnorm) {
        d3 n = crs(sub(b, a), sub(c, a));
        st3(tnorm + 3 * (size_t)j, scl(1.0 / sqrt(dot(n, n)), n));
    }
    if (tgrad) {
        const d3 ab = sub(b, a), bc = sub(c, b);
        d3 n = crs(ab, bc);
        n = scl(1.0 / dot(n, n), n);
        st3(tgrad + 3 * (size_t)j, add(scl(pot[ic] - pot[ib], crs(n, ab)), scl(pot[ia] - pot[ib], crs(n, bc))));
    }
    if (tvol) tvol[j] = dot(crs(a, b), c) / 6.0;
}

// generate_vertex_normals with Max 1999 weights (Mesh.cpp:141-163)
__global__ __launch_bounds__(256) void k_vertex_normals(int N, const double* __restrict__ xyz,
                                                        const uint32_t* __restrict__ tri,
                                                        const int32_t* __restrict__ vptr,
                                                        const int32_t* __restrict__ vtri, double* __restrict__ vn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    d3 acc = {0, 0, 0};
    for (int e = vptr[i]; e < vptr[i + 1]; ++e) {
        const uint32_t* t = tri + 3 * (size_t)vtri[e];
        const int s = (t[0] == (uint32_t)i) ? 0 : ((t[1] == (uint32_t)i) ? 1 : 2);
        const d3 a = ld3(xyz + 3 * (size_t)t[s]), b = ld3(xyz + 3 * (size_t)t[(s + 1) % 3]),
                 c = ld3(xyz + 3 * (size_t)t[(s + 2) % 3]);
        const d3 B = sub(b, a), C = sub(c, a);
        acc = add(acc, scl(1.0 / (dot(B, B) * dot(C, C)), crs(B, C)));
    }
    st3(vn + 3 * (size_t)i, scl(1.0 / sqrt(dot(acc, acc)), acc));
}

// Eigen::FullPivLU<Matrix<double,6,3>> solve: complete pivoting (first maximum
// in column-major order), solution of the three pivot equations.
__device__ inline void fullpiv_6x3(double Mx[6][3], double bv[6], double x[3]) {
    int rowt[3], colt[3], perm[3] = {0, 1, 2};
    int rank = 3;
    double maxpivot = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= rank) break;
        int br = k, bc = k;
        double best = -1.0;
        for (int j = k; j < 3; ++j)
            for (int i = k; i < 6; ++i) {
                const double v = fabs(Mx[i][j]);
                if (v > best) { best = v; br = i; bc = j; }
            }
        if (best == 0.0) {
            rank = k;
            for (int i = k; i < 3; ++i) { rowt[i] = i; colt[i] = i; }
            break;
        }
        maxpivot = fmax(maxpivot, best);
        rowt[k] = br; colt[k] = bc;
        if (br != k)
            for (int j = 0; j < 3; ++j) { const double t = Mx[k][j]; Mx[k][j] = Mx[br][j]; Mx[br][j] = t; }
        if (bc != k)
            for (int i = 0; i < 6; ++i) { const double t = Mx[i][k]; Mx[i][k] = Mx[i][bc]; Mx[i][bc] = t; }
        for (int i = k + 1; i < 6; ++i) Mx[i][k] /= Mx[k][k];
        if (k < 2)
            for (int j = k + 1; j < 3; ++j)
                for (int i = k + 1; i < 6; ++i) Mx[i][j] -= Mx[i][k] * Mx[k][j];
    }
    int nz = 0;
    for (int k = 0; k < rank; ++k)
        if (fabs(Mx[k][k]) > 2.220446049250313e-16 * 3.0 * maxpivot) ++nz;
    for (int k = 0; k < 3; ++k) { const int t = perm[k]; perm[k] = perm[colt[k]]; perm[colt[k]] = t; }
    double c[6];
    for (int i = 0; i < 6; ++i) c[i] = bv[i];
    for (int k = 0; k < 3; ++k) { const double t = c[k]; c[k] = c[rowt[k]]; c[rowt[k]] = t; }
    for (int i = 1; i < 3; ++i)
        for (int j = 0; j < i; ++j) c[i] -= Mx[i][j] * c[j];
    for (int i = nz - 1; i >= 0; --i) {
        for (int j = i + 1; j < nz; ++j) c[i] -= Mx[i][j] * c[j];
        c[i] /= Mx[i][i];
    }
    x[0] = x[1] = x[2] = 0.0;
    for (int i = 0; i < nz; ++i) x[perm[i]] = c[i];
}

// curvatures (Mesh.cpp:256-319): mean curvature per triangle
__global__ __launch_bounds__(256) void k_tri_curv(int M, const double* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                  const double* __restrict__ vn, double* __restrict__ tk) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t ia = tri[3 * j], ib = tri[3 * j + 1], ic = tri[3 * j + 2];
    const d3 a = ld3(xyz + 3 * (size_t)ia), b = ld3(xyz + 3 * (size_t)ib), c = ld3(xyz + 3 * (size_t)ic);
    const d3 e0 = sub(b, a), e1 = sub(c, b), e2 = sub(a, c);
    const d3 u = scl(1.0 / sqrt(dot(e0, e0)), e0);
    d3 nr = crs(e0, e1);
    nr = scl(1.0 / sqrt(dot(nr, nr)), nr);
    const d3 w = crs(nr, u);
    double Mx[6][3] = {{dot(e0, u), dot(e0, w), 0.0}, {0.0, dot(e0, u), dot(e0, w)},
                       {dot(e1, u), dot(e1, w), 0.0}, {0.0, dot(e1, u), dot(e1, w)},
                       {dot(e2, u), dot(e2, w), 0.0}, {0.0, dot(e2, u), dot(e2, w)}};
    const d3 na = ld3(vn + 3 * (size_t)ia), nb = ld3(vn + 3 * (size_t)ib), nc = ld3(vn + 3 * (size_t)ic);
    double bv[6] = {dot(sub(nb, na), u), dot(sub(nb, na), w), dot(sub(nc, nb), u),
                    dot(sub(nc, nb), w), dot(sub(na, nc), u), dot(sub(na, nc), w)};
    double x[3];
    fullpiv_6x3(Mx, bv, x);
    tk[j] = 0.5 * (x[0] + x[2]);
}

// one work-group per loose part, fixed-order tree reduction
__global__ __launch_bounds__(256) void k_part_volume(const int32_t* __restrict__ tptr, const int32_t* __restrict__ ptri,
                                                     const double* __restrict__ tvol, double* __restrict__ pvol) {
    __shared__ double s[256];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int e = tptr[c] + threadIdx.x; e < tptr[c + 1]; e += 256) acc += tvol[ptri[e]];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) pvol[c] = s[0];
}

// LinLinSim.cpp:48-59: v_i = mean over T of (grad_T phi + psi_i n_T)
__global__ __launch_bounds__(256) void k_vertex_velocity(int N, const int32_t* __restrict__ vptr,
                                                         const int32_t* __restrict__ vtri,
                                                         const double* __restrict__ tnorm,
                                                         const double* __restrict__ tgrad,
                                                         const double* __restrict__ psi, double* __restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double ps = psi[i];
    d3 g = {0, 0, 0};
    double num = 0.0;
    for (int e = vptr[i]; e < vptr[i + 1]; ++e) {
        const int t = vtri[e];
        g = add(g, add(ld3(tgrad + 3 * (size_t)t), scl(ps, ld3(tnorm + 3 * (size_t)t))));
        num += 1.0;
    }
    st3(vel + 3 * (size_t)i, scl(1.0 / num, g));
}

// potential_t (Simulation.cpp:73-75) per vertex with its part's volume and
// kappa averaged over the incident triangles (LinLinSim.cpp:122-146)
__global__ __launch_bounds__(256) void k_vertex_dpot(int N, StageParams sp, double t, const double* __restrict__ xyz,
                                                     const double* __restrict__ vel, const int32_t* __restrict__ vptr,
                                                     const int32_t* __restrict__ vtri, const double* __restrict__ tk,
                                                     const int32_t* __restrict__ vpart,
                                                     const double* __restrict__ pvol,
                                                     const double* __restrict__ pphys, double* __restrict__ dpot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double kap = 0.0;
    if (tk) {
        double w = 0.0;
        for (int e = vptr[i]; e < vptr[i + 1]; ++e) { kap += tk[vtri[e]]; w += 1.0; }
        kap /= w;
    }
    const d3 v = ld3(vel + 3 * (size_t)i);
    const int c = vpart[i];
    const double* pp = pphys + 4 * (size_t)c;   // part c: V_0, amplitude, frequency, phase
    const double field = pp[1] * sin(2.0 * PI * pp[2] * t + pp[3]);
    dpot[i] = 2.0 * sp.sigma * kap + 0.5 * dot(v, v) + sp.p_inf - sp.epsilon * pow(pp[0] / pvol[c], sp.gamma) +
              field;
}

// stage state: xs = c kx + x1, ps = c kp + p1 (LinLinSim.cpp:218-232)
__global__ __launch_bounds__(256) void k_axpy_state(int N, double c, const double* __restrict__ kx,
                                                    const double* __restrict__ x1, double* __restrict__ xs,
                                                    const double* __restrict__ kp, const double* __restrict__ p1,
                                                    double* __restrict__ ps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) xs[3 * (size_t)i + d] = c * kx[3 * (size_t)i + d] + x1[3 * (size_t)i + d];
    ps[i] = c * kp[i] + p1[i];
}

// x (N x 3) = c dx (+ x when acc), p (N) = c dp (+ p when acc)
__global__ __launch_bounds__(256) void k_axpy_pair(int N, double c, const double* dx, double* x, const double* dp,
                                                   double* p, int acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t e = 3 * (size_t)i + d;
        x[e] = c * dx[e] + (acc ? x[e] : 0.0);
    }
    p[i] = c * dp[i] + (acc ? p[i] : 0.0);
}

// RK4 combination (LinLinSim.cpp:236-240) incl. the wall guard nopenetration
// (LinLinSim.cpp:162-190); kx, kp hold the 4 stages back to back.
__global__ __launch_bounds__(256) void k_rk4_final(int N, double dt, const double* __restrict__ kx,
                                                   const double* __restrict__ kp, const double* __restrict__ x1,
                                                   double* __restrict__ x, double* __restrict__ phi,
                                                   double* __restrict__ avg, int guards, double eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t n3 = 3 * (size_t)N;
    double a[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t e = 3 * (size_t)i + d;
        a[d] = (((kx[e] + 2.0 * kx[n3 + e]) + 2.0 * kx[2 * n3 + e]) + kx[3 * n3 + e]) * (1.0 / 6.0);
        x[e] = dt * a[d] + x1[e];
    }
    phi[i] = (dt / 6.0) * (((kp[i] + 2.0 * kp[N + i]) + 2.0 * kp[2 * (size_t)N + i]) + kp[3 * (size_t)N + i]) + phi[i];
    if (guards) {
        const double cx = a[0];
        if (cx * dt < 0.0) {
            bool flag = false;
            const double xpos = x1[3 * (size_t)i];
            double xf = x[3 * (size_t)i];
            if (xpos < eps) {
                xf = xpos * exp(cx / xpos * dt);
                flag = true;
            } else {
                const double te = (eps - xpos) / cx;
                if (fabs(te) < fabs(dt)) { xf = eps * exp(cx / eps * (dt - te)); flag = true; }
            }
            if (flag) { x[3 * (size_t)i] = xf; a[0] = (xf - xpos) / dt; }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) avg[3 * (size_t)i + d] = a[d];
}

__global__ __launch_bounds__(256) void k_psi_project(int N, const double* __restrict__ avg, const double* __restrict__ vn,
                                                     double* __restrict__ psi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) psi[i] = dot(ld3(avg + 3 * (size_t)i), ld3(vn + 3 * (size_t)i));
}

__global__ __launch_bounds__(256) void k_euler(int N, double dt, const double* __restrict__ vel,
                                               const double* __restrict__ dpot, double* __restrict__ x,
                                               double* __restrict__ phi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) x[3 * (size_t)i + d] = vel[3 * (size_t)i + d] * dt + x[3 * (size_t)i + d];
    phi[i] = dt * dpot[i] + phi[i];
}

// out[0] = max |dpot|, out[1] = max |v|^2, out[2] = 1 if some x < 0
__global__ __launch_bounds__(256) void k_maxes(int N, const double* __restrict__ vel, const double* __restrict__ dpot,
                                               const double* __restrict__ xyz, double* __restrict__ out) {
    __shared__ double s0[256], s1[256], s2[256];
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) {
        if (dpot) m0 = fmax(m0, fabs(dpot[i]));
        if (vel) { const d3 v = ld3(vel + 3 * (size_t)i); m1 = fmax(m1, dot(v, v)); }
        if (xyz && xyz[3 * (size_t)i] < 0.0) m2 = 1.0;
    }
    s0[threadIdx.x] = m0; s1[threadIdx.x] = m1; s2[threadIdx.x] = m2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; out[2] = s2[0]; }
}

// ---- pinned contact line (ColocSimPin) ----
// The last N - first vertices are pinned on the wall x = 0.

// dst[i] = src[i] for the pinned vertices
__global__ __launch_bounds__(256) void k_pin_copy_tail(int N, int first, const double* __restrict__ src,
                                                       double* __restrict__ dst) {
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = src[i];
}

// pz = pot with the pinned entries zeroed: the fused H pz is then the sum over
// the free columns only
__global__ __launch_bounds__(256) void k_pin_zero_tail(int N, int first, const double* __restrict__ pot,
                                                       double* __restrict__ pz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) pz[i] = i < first ? pot[i] : 0.0;
}

// ColocSimPin::assemble_matrices (ColocSimPin.cpp:113-117) applied to the
// system G x = H pot: pinned column j of G becomes -H(:, j) (also in the FP32
// copy) and, with H(:, j) = -G_old(:, j), rhs_i gains -sum_j G_old(i, j) pot_j.
// One thread per row; the column loop reads consecutive rows (coalesced).
__global__ __launch_bounds__(256) void k_pin_swap(int N, int first, double* __restrict__ G, float* __restrict__ G32,
                                                  const double* __restrict__ H, const double* __restrict__ pot,
                                                  double* __restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double acc = rhs[i];
    for (int j = first; j < N; ++j) {
        const size_t e = (size_t)j * N + i;
        const double g = G[e], h = H[e];
        G[e] = -h;
        if (G32) G32[e] = (float)(-h);
        acc += -g * pot[j];
    }
    rhs[i] = acc;
}

// ColocSimPin::position_t (ColocSimPin.cpp:176-188): a pinned vertex moves
// along the wall, in the direction of its normal with the x part removed, at
// the drift velocity of cos(theta) = n.x (ColocSimPin.cpp:126-136:
// beta (cos_eq - c) + c_nl c^3 / (c^2 - 1)); its potential entry takes the
// solved value.
__global__ __launch_bounds__(256) void k_pin_drift(int N, int first, double beta, double cos_eq, double c_nl,
                                                   const double* __restrict__ vn, const double* __restrict__ psil,
                                                   double* __restrict__ vel, double* __restrict__ pot) {
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    d3 n = ld3(vn + 3 * (size_t)i);
    const double c = n.x;
    n.x = 0.0;
    n = scl(1.0 / sqrt(dot(n, n)), n);
    const double drift = beta * (cos_eq - c) + c_nl * (c * c * c) / (c * c - 1.0);
    n = scl(drift, n);
    vel[3 * (size_t)i] = n.x;
    vel[3 * (size_t)i + 1] = n.y;
    vel[3 * (size_t)i + 2] = n.z;
    pot[i] = psil[i];
}

static inline int nb(int64_t n) { return (int)((n + 255) / 256); }

void pin_copy_tail(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* src, double* dst) {
    if (npin <= 0) return;
    k_pin_copy_tail<<<nb(npin), 256, 0, ctx->stream>>>((int)N, (int)(N - npin), src, dst);
    BH_HIP(hipGetLastError());
}
void pin_zero_tail(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* pot, double* pz) {
    k_pin_zero_tail<<<nb(N), 256, 0, ctx->stream>>>((int)N, (int)(N - npin), pot, pz);
    BH_HIP(hipGetLastError());
}
void pin_swap_columns(bemhip_ctx* ctx, int64_t N, int64_t npin, double* G, float* G32, const double* H,
                      const double* pot, double* rhs) {
    if (npin <= 0) return;
    k_pin_swap<<<nb(N), 256, 0, ctx->stream>>>((int)N, (int)(N - npin), G, G32, H, pot, rhs);
    BH_HIP(hipGetLastError());
}
void pin_drift(bemhip_ctx* ctx, int64_t N, int64_t npin, const double* vn, const double* psil, double* vel,
               double* pot) {
    if (npin <= 0) return;
    // the reference's constants (ColocSimPin.cpp:126-130); cos(130 deg) on the host
    const double beta = 0.2, cos_eq = std::cos(130.0 / 180.0 * PI), c_nl = 0.01;
    k_pin_drift<<<nb(npin), 256, 0, ctx->stream>>>((int)N, (int)(N - npin), beta, cos_eq, c_nl, vn, psil, vel, pot);
    BH_HIP(hipGetLastError());
}

void tri_geometry(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* pot, double* tnorm,
                  double* tgrad, double* tvol) {
    k_tri_geom<<<nb(topo.M), 256, 0, ctx->stream>>>((int)topo.M, xyz, topo.tri.as<uint32_t>(), pot, tnorm, tgrad, tvol);
    BH_HIP(hipGetLastError());
}
void vertex_normals(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, double* vn) {
    k_vertex_normals<<<nb(topo.N), 256, 0, ctx->stream>>>((int)topo.N, xyz, topo.tri.as<uint32_t>(),
                                                          topo.vptr.as<int32_t>(), topo.vtri.as<int32_t>(), vn);
    BH_HIP(hipGetLastError());
}
void tri_curvature(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* vn, double* tk) {
    k_tri_curv<<<nb(topo.M), 256, 0, ctx->stream>>>((int)topo.M, xyz, topo.tri.as<uint32_t>(), vn, tk);
    BH_HIP(hipGetLastError());
}
void part_volumes(bemhip_ctx* ctx, const DevTopology& topo, const double* tvol, double* pvol) {
    k_part_volume<<<topo.nparts, 256, 0, ctx->stream>>>(topo.part_tptr.as<int32_t>(), topo.part_tri.as<int32_t>(),
                                                        tvol, pvol);
    BH_HIP(hipGetLastError());
}
void vertex_velocity(bemhip_ctx* ctx, const DevTopology& topo, const double* tnorm, const double* tgrad,
                     const double* psi, double* vel) {
    k_vertex_velocity<<<nb(topo.N), 256, 0, ctx->stream>>>((int)topo.N, topo.vptr.as<int32_t>(),
                                                           topo.vtri.as<int32_t>(), tnorm, tgrad, psi, vel);
    BH_HIP(hipGetLastError());
}
void vertex_dpot(bemhip_ctx* ctx, const DevTopology& topo, const StageParams& sp, const double* xyz,
                 const double* vel, const double* tk, const double* pvol, const double* pphys, double t,
                 double* dpot) {
    k_vertex_dpot<<<nb(topo.N), 256, 0, ctx->stream>>>((int)topo.N, sp, t, xyz, vel, topo.vptr.as<int32_t>(),
                                                       topo.vtri.as<int32_t>(), tk, topo.vpart.as<int32_t>(), pvol,
                                                       pphys, dpot);
    BH_HIP(hipGetLastError());
}
void axpy_state(bemhip_ctx* ctx, int64_t N, double c, const double* kx, const double* x1, double* xs,
                const double* kp, const double* p1, double* ps) {
    k_axpy_state<<<nb(N), 256, 0, ctx->stream>>>((int)N, c, kx, x1, xs, kp, p1, ps);
    BH_HIP(hipGetLastError());
}
void axpy_pair(bemhip_ctx* ctx, int64_t N, double c, const double* dx, double* x, const double* dp, double* p,
               bool accumulate) {
    k_axpy_pair<<<nb(N), 256, 0, ctx->stream>>>((int)N, c, dx, x, dp, p, accumulate ? 1 : 0);
    BH_HIP(hipGetLastError());
}
void rk4_final(bemhip_ctx* ctx, int64_t N, double dt, const double* kx, const double* kp, const double* x1,
               double* x, double* phi, double* avg, int wall_guards, double wall_eps) {
    k_rk4_final<<<nb(N), 256, 0, ctx->stream>>>((int)N, dt, kx, kp, x1, x, phi, avg, wall_guards, wall_eps);
    BH_HIP(hipGetLastError());
}
void psi_project(bemhip_ctx* ctx, int64_t N, const double* avg, const double* vn, double* psi) {
    k_psi_project<<<nb(N), 256, 0, ctx->stream>>>((int)N, avg, vn, psi);
    BH_HIP(hipGetLastError());
}
void euler_update(bemhip_ctx* ctx, int64_t N, double dt, const double* vel, const double* dpot, double* x,
                  double* phi) {
    k_euler<<<nb(N), 256, 0, ctx->stream>>>((int)N, dt, vel, dpot, x, phi);
    BH_HIP(hipGetLastError());
}
void dt_reduce(bemhip_ctx* ctx, int64_t N, const double* vel, const double* dpot, double* out2) {
    double* d = static_cast<double*>(ctx->scal.ensure(sizeof(double) * 1024)) + 1000;
    k_maxes<<<1, 256, 0, ctx->stream>>>((int)N, vel, dpot, nullptr, d);
    BH_HIP(hipMemcpyAsync(out2, d, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
}
int any_negative_x(bemhip_ctx* ctx, int64_t N, const double* xyz) {
    double* d = static_cast<double*>(ctx->scal.ensure(sizeof(double) * 1024)) + 1000;
    k_maxes<<<1, 256, 0, ctx->stream>>>((int)N, nullptr, nullptr, xyz, d);
    double h[3];
    BH_HIP(hipMemcpyAsync(h, d, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return h[2] > 0.0;
}

}  // namespace bemhip
