// Reverse-mode (adjoint) kernels of one RK4 stage: given the cotangents of a
// stage's outputs (vertex velocities kx = position_t and Dphi/Dt kp =
// pot_t_multi, LinLinSim.cpp:26-120), accumulate the cotangents of the stage's
// inputs (positions xs, potential ps).  The reference has no derivative code;
// this is the transposed linearisation of the forward stage as computed by
// assemble.hip / dense.hip / stage.hip, for sigma = 0 and no wall image.
//
// With psi = G^-1 H phi the psi-cotangent psibar enters through
// lambda = G^-T psibar (transposed solve with the stage's stored LU factors),
// after which
//     phibar += H^T lambda,
//     xbar   += d/dx [ lambda^T (H phi - G psi) ]   (psi, phi, lambda held fixed),
// evaluated by re-running the panel quadrature (no dG/dx is formed): per pair
// (row i, triangle T not containing i), with N = ab x bc, J = |N|,
// z_q = y_q - x_i, r_q = |z_q|, A_q / P_q the linear interpolants of psi / phi,
//     S = -lambda_i sum_q w_q [ J A_q / r_q + (P_q - phi_i) (z_q . N) / r_q^3 ],
// and the self panels (i in T) contribute -lambda_i sum_k G_k(i,T) psi_k.
// Every sum is split into per-triangle and per-row partial arrays that a
// final per-vertex gather adds in a fixed order: no atomics, deterministic.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "quad_rules.hpp"
#include "vec3.hpp"

namespace bemhip {

namespace {

__device__ inline d3 shfl_sum3(d3 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
        v.z += __shfl_xor(v.z, off);
    }
    return v;
}
__device__ inline double shfl_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// corner cotangents of N = ab x bc, ab = b - a, bc = c - b, for N-cotangent Nb
__device__ inline void normal_to_corners(d3 ab, d3 bc, d3 Nb, d3& ga, d3& gb, d3& gc) {
    const d3 abb = crs(bc, Nb), bcb = crs(Nb, ab);
    ga = sub(ga, abb);
    gb = add(gb, sub(abb, bcb));
    gc = add(gc, bcb);
}

}  // namespace

// kxbar += kpbar * v  (d/dv of |v|^2 / 2)
__global__ __launch_bounds__(256) void k_adj_kx(int N, const double* __restrict__ vel, const double* __restrict__ kpbar,
                                                double* __restrict__ kxbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double p = kpbar[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) kxbar[3 * (size_t)i + d] += p * vel[3 * (size_t)i + d];
}

// Volume cotangent of every loose part for the Bernoulli term
// -eps (V_0 / V_c)^gamma: coef_c = eps gamma (V_0/V_c)^gamma / V_c * sum_{i in c} kpbar_i.
// One work-group per part, fixed summation order.
__global__ __launch_bounds__(256) void k_adj_part_coef(const int32_t* __restrict__ part_vptr,
                                                       const int32_t* __restrict__ part_vtx,
                                                       const double* __restrict__ kpbar,
                                                       const double* __restrict__ pvol,
                                                       const double* __restrict__ pphys, double eps, double gam,
                                                       double* __restrict__ coef) {
    __shared__ double red[256];
    const int c = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int e = part_vptr[c] + tid; e < part_vptr[c + 1]; e += 256) s += kpbar[part_vtx[e]];
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) {
        const double V = pvol[c], V0 = pphys[4 * (size_t)c];
        coef[c] = red[0] * eps * gam * pow(V0 / V, gam) / V;
    }
}

// Per triangle: cotangents of its three corners (9) and of the potential at
// them (3) from the vertex-velocity average v_i = mean_{T ni i}(g_T + psi_i n_T)
// (LinLinSim.cpp:48-59; g_T the Dombre tangential gradient, LinLinSim.cpp:288-305)
// and from the part volumes (coefficient vcoef of the triangle's part).
// psibar is produced separately by k_adj_psibar.
__global__ __launch_bounds__(256) void k_adj_tri(int M, const double* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                 const double* __restrict__ pot, const double* __restrict__ psi,
                                                 const double* __restrict__ kxbar, const int32_t* __restrict__ vptr,
                                                 const int32_t* __restrict__ vpart, const double* __restrict__ vcoef,
                                                 double* __restrict__ tc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t ia = tri[3 * (size_t)j], ib = tri[3 * (size_t)j + 1], ic = tri[3 * (size_t)j + 2];
    const d3 a = ld3(xyz + 3 * (size_t)ia), b = ld3(xyz + 3 * (size_t)ib), c = ld3(xyz + 3 * (size_t)ic);
    const double pa = pot[ia], pb = pot[ib], pc = pot[ic];
    const d3 ua = scl(1.0 / (double)(vptr[ia + 1] - vptr[ia]), ld3(kxbar + 3 * (size_t)ia));
    const d3 ub = scl(1.0 / (double)(vptr[ib + 1] - vptr[ib]), ld3(kxbar + 3 * (size_t)ib));
    const d3 uc = scl(1.0 / (double)(vptr[ic + 1] - vptr[ic]), ld3(kxbar + 3 * (size_t)ic));
    const d3 gbar = add(add(ua, ub), uc);
    const d3 nbar = add(add(scl(psi[ia], ua), scl(psi[ib], ub)), scl(psi[ic], uc));
    const d3 ab = sub(b, a), bc = sub(c, b);
    const d3 N = crs(ab, bc);
    const double NN = dot(N, N), Jn = sqrt(NN);
    const d3 m = scl(1.0 / NN, N), nh = scl(1.0 / Jn, N);
    // g = m x w, w = (pc - pb) ab + (pa - pb) bc
    const d3 w = add(scl(pc - pb, ab), scl(pa - pb, bc));
    const double pbar_c = dot(gbar, crs(m, ab)), pbar_a = dot(gbar, crs(m, bc));
    const double pbar_b = -(pbar_a + pbar_c);
    const d3 mbar = crs(w, gbar), wbar = crs(gbar, m);
    d3 abb = scl(pc - pb, wbar), bcb = scl(pa - pb, wbar);
    // m = N / |N|^2 and n = N / |N|
    d3 Nb = sub(scl(1.0 / NN, mbar), scl(2.0 * dot(N, mbar) / (NN * NN), N));
    Nb = add(Nb, scl(1.0 / Jn, sub(nbar, scl(dot(nh, nbar), nh))));
    abb = add(abb, crs(bc, Nb));
    bcb = add(bcb, crs(Nb, ab));
    d3 ga = scl(-1.0, abb), gb = sub(abb, bcb), gc = bcb;
    // part volume: V_T = a . (b x c) / 6 (Mesh.cpp:206-218)
    const double cf = vcoef[vpart[ia]] * (1.0 / 6.0);
    ga = add(ga, scl(cf, crs(b, c)));
    gb = add(gb, scl(cf, crs(c, a)));
    gc = add(gc, scl(cf, crs(a, b)));
    double* o = tc + 12 * (size_t)j;
    st3(o, ga); st3(o + 3, gb); st3(o + 6, gc);
    o[9] = pbar_a; o[10] = pbar_b; o[11] = pbar_c;
}

// psibar_i = (1/deg_i) sum_{T ni i} kxbar_i . n_T
__global__ __launch_bounds__(256) void k_adj_psibar(int N, const double* __restrict__ xyz,
                                                    const uint32_t* __restrict__ tri, const int32_t* __restrict__ vptr,
                                                    const int32_t* __restrict__ vtri, const double* __restrict__ kxbar,
                                                    double* __restrict__ psibar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const d3 u = ld3(kxbar + 3 * (size_t)i);
    double s = 0.0;
    for (int e = vptr[i]; e < vptr[i + 1]; ++e) {
        const uint32_t* t = tri + 3 * (size_t)vtri[e];
        const d3 a = ld3(xyz + 3 * (size_t)t[0]), b = ld3(xyz + 3 * (size_t)t[1]), c = ld3(xyz + 3 * (size_t)t[2]);
        const d3 n = crs(sub(b, a), sub(c, a));
        s += dot(u, n) / sqrt(dot(n, n));
    }
    psibar[i] = s / (double)(vptr[i + 1] - vptr[i]);
}

// Far-field backward assembly.  Lane = triangle (256 per work-group), the
// work-group walks a chunk of collocation rows.  Per triangle the lane keeps
// its corner / potential / normal cotangents; per row the four row cotangents
// (x_i, phi_i) are wave-reduced, staged in LDS for ADJ_RB rows and written as
// this triangle block's partial.
constexpr int ADJ_RB = 32;

__global__ __launch_bounds__(256) void k_adj_far(int N, int M, int rows_per_chunk, const double* __restrict__ xyz,
                                                 const uint32_t* __restrict__ tri, const double* __restrict__ pot,
                                                 const double* __restrict__ psi, const double* __restrict__ lam,
                                                 double* __restrict__ tpart, double* __restrict__ rpart) {
    __shared__ double red[4][ADJ_RB][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int j = blockIdx.x * 256 + tid;
    const bool tv = j < M;
    const int jj = tv ? j : M - 1;
    const uint32_t ia = tri[3 * (size_t)jj], ib = tri[3 * (size_t)jj + 1], ic = tri[3 * (size_t)jj + 2];
    const d3 a = ld3(xyz + 3 * (size_t)ia), b = ld3(xyz + 3 * (size_t)ib), c = ld3(xyz + 3 * (size_t)ic);
    const d3 ab = sub(b, a), bc = sub(c, b);
    const d3 Nv = crs(ab, bc);
    const double J = sqrt(dot(Nv, Nv));
    const double sa = psi[ia], sb = psi[ib], sc = psi[ic];
    const double fa = pot[ia], fb = pot[ib], fc = pot[ic];
    d3 g0 = {0, 0, 0}, g1 = {0, 0, 0}, g2 = {0, 0, 0}, Nb = {0, 0, 0};
    double Jb = 0.0, q0 = 0.0, q1 = 0.0, q2 = 0.0;

    const int r0 = blockIdx.y * rows_per_chunk, r1 = min(N, r0 + rows_per_chunk);
    for (int rb = r0; rb < r1; rb += ADJ_RB) {
        const int nr = min(ADJ_RB, r1 - rb);
        for (int k = 0; k < nr; ++k) {
            const int i = rb + k;
            const d3 x = ld3(xyz + 3 * (size_t)i);
            const double li = lam[i], fi = pot[i];
            d3 gx = {0, 0, 0};
            double gp = 0.0;
            const bool far = tv && (uint32_t)i != ia && (uint32_t)i != ib && (uint32_t)i != ic;
            if (far) {
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    const double u = TQ_X[q] + TQ_Y[q], v = TQ_Y[q];
                    const double b0 = 1.0 - u, b1 = u - v, b2 = v;   // basis at the point
                    const d3 z = sub(add(a, add(scl(u, ab), scl(v, bc))), x);
                    const double ir = rsqrt_nr(dot(z, z));
                    const double ir2 = ir * ir, ir3 = ir * ir2;
                    const double zn = dot(z, Nv);
                    const double Aq = b0 * sa + b1 * sb + b2 * sc;
                    const double Bq = b0 * fa + b1 * fb + b2 * fc - fi;
                    const double wl = -li * TQ_W[q];
                    const double al = wl * (-J * Aq * ir3 - 3.0 * Bq * zn * ir3 * ir2);
                    const double be = wl * Bq * ir3;
                    const d3 dz = add(scl(al, z), scl(be, Nv));
                    gx = sub(gx, dz);
                    g0 = fma3(b0, dz, g0);
                    g1 = fma3(b1, dz, g1);
                    g2 = fma3(b2, dz, g2);
                    Jb += wl * Aq * ir;
                    Nb = fma3(be, z, Nb);
                    const double h = wl * zn * ir3;
                    q0 += b0 * h; q1 += b1 * h; q2 += b2 * h;
                    gp -= h;
                }
            }
            gx = shfl_sum3(gx);
            gp = shfl_sum(gp);
            if (lane == 0) {
                red[wv][k][0] = gx.x; red[wv][k][1] = gx.y; red[wv][k][2] = gx.z; red[wv][k][3] = gp;
            }
        }
        __syncthreads();
        if (tid < 4 * nr) {
            const int k = tid >> 2, cc = tid & 3;
            const double s = ((red[0][k][cc] + red[1][k][cc]) + red[2][k][cc]) + red[3][k][cc];
            rpart[((size_t)blockIdx.x * N + rb + k) * 4 + cc] = s;
        }
        __syncthreads();
    }
    if (tv) {
        // J = |N|: its cotangent joins N's, then N = ab x bc goes to the corners
        const d3 Nt = add(Nb, scl(Jb / J, Nv));
        normal_to_corners(ab, bc, Nt, g0, g1, g2);
        double* o = tpart + ((size_t)blockIdx.y * M + j) * 12;
        st3(o, g0); st3(o + 3, g1); st3(o + 6, g2);
        o[9] = q0; o[10] = q1; o[11] = q2;
    }
}

// Self panels: one thread per (triangle, corner k); the collocation point is
// corner k and the polar rule of integrate_identical_coloc (Integrator.cpp:12-41)
// gives G_k; S = -lambda_k sum_k' G_k' psi_k'.  Output: the nine corner
// cotangents in the triangle's own corner order, sc[(3 T + k) * 9 ..].
__global__ __launch_bounds__(256) void k_adj_self(int M, const double* __restrict__ xyz,
                                                  const uint32_t* __restrict__ tri, const double* __restrict__ psi,
                                                  const double* __restrict__ lam, double* __restrict__ sc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * M) return;
    const int j = e / 3, k = e % 3;
    const uint32_t* t = tri + 3 * (size_t)j;
    const uint32_t i0 = t[k], i1 = t[(k + 1) % 3], i2 = t[(k + 2) % 3];   // cyclic_reorder (Triplet.hpp:55-72)
    const d3 a = ld3(xyz + 3 * (size_t)i0), b = ld3(xyz + 3 * (size_t)i1), c = ld3(xyz + 3 * (size_t)i2);
    const d3 ab = sub(b, a), bc = sub(c, b);
    const d3 Nv = crs(ab, bc);
    const double J = sqrt(dot(Nv, Nv));
    const double jac = 0.5 * J * (PI / 4.0);
    const double li = lam[i0], s0 = psi[i0], s1 = psi[i1], s2 = psi[i2];
    double S = 0.0;
    d3 abb = {0, 0, 0}, bcb = {0, 0, 0};
#pragma unroll
    for (int p = 0; p < 7; ++p) {
        const double th = GQ_X[p] * (PI / 4.0);
        const double cs = cos(th), sn = sin(th);
        const d3 d = add(scl(cs, ab), scl(sn, bc));
        const double dist = sqrt(dot(d, d));
        const double F = cs * s0 + (cs - sn) * s1 + sn * s2;
        const double K = -li * jac * GQ_W[p] * F / (cs * cs);
        S += K / dist;
        const d3 dd = scl(-K / (dist * dist * dist), d);   // dS/dd
        abb = fma3(cs, dd, abb);
        bcb = fma3(sn, dd, bcb);
    }
    // S is proportional to J = |N|
    const d3 Nb = scl(S / (J * J), Nv);
    abb = add(abb, crs(bc, Nb));
    bcb = add(bcb, crs(Nb, ab));
    const d3 g[3] = {scl(-1.0, abb), sub(abb, bcb), bcb};   // reordered corners 0, 1, 2
    double* o = sc + 9 * (size_t)e;
#pragma unroll
    for (int r = 0; r < 3; ++r) st3(o + 3 * ((k + r) % 3), g[r]);
}

// Per vertex: sum every partial in a fixed order.
__global__ __launch_bounds__(256) void k_adj_gather(int N, int M, const uint32_t* __restrict__ tri,
                                                    const int32_t* __restrict__ vptr, const int32_t* __restrict__ vtri,
                                                    const double* __restrict__ tc, const double* __restrict__ tpart,
                                                    int nchunks, const double* __restrict__ rpart, int ntb,
                                                    const double* __restrict__ sc, const double* __restrict__ lam,
                                                    double* __restrict__ xbar, double* __restrict__ pbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    d3 g = {0, 0, 0};
    double p = 0.0;
    for (int e = vptr[i]; e < vptr[i + 1]; ++e) {
        const int j = vtri[e];
        const uint32_t* t = tri + 3 * (size_t)j;
        const int s = (t[0] == (uint32_t)i) ? 0 : ((t[1] == (uint32_t)i) ? 1 : 2);
        if (tc) {
            g = add(g, ld3(tc + 12 * (size_t)j + 3 * s));
            p += tc[12 * (size_t)j + 9 + s];
        }
        if (tpart)
            for (int ch = 0; ch < nchunks; ++ch) {
                const double* o = tpart + ((size_t)ch * M + j) * 12;
                g = add(g, ld3(o + 3 * s));
                p += o[9 + s];
            }
        if (sc)
            for (int k = 0; k < 3; ++k) g = add(g, ld3(sc + 9 * (3 * (size_t)j + k) + 3 * s));
    }
    if (rpart)
        for (int b = 0; b < ntb; ++b) {
            const double* o = rpart + ((size_t)b * N + i) * 4;
            g = add(g, ld3(o));
            p += o[3];
        }
    if (lam) p -= 4.0 * PI * lam[i];   // the -4 pi phi_i of the closed H row
    xbar[3 * (size_t)i] += g.x;
    xbar[3 * (size_t)i + 1] += g.y;
    xbar[3 * (size_t)i + 2] += g.z;
    if (pbar) pbar[i] += p;
}

// ------------------------------------------------------------ host side --
static int nb256(int64_t n) { return (int)((n + 255) / 256); }

void adj_stage_pre(bemhip_ctx* ctx, const DevTopology& topo, const AdjStageIn& in, double* kxbar, const double* kpbar,
                   double* psibar) {
    hipStream_t st = ctx->stream;
    const int N = (int)topo.N, M = (int)topo.M;
    k_adj_kx<<<nb256(N), 256, 0, st>>>(N, in.vel, kpbar, kxbar);
    double* coef = static_cast<double*>(ctx->adj_coef.ensure(sizeof(double) * std::max(1, topo.nparts)));
    k_adj_part_coef<<<topo.nparts, 256, 0, st>>>(topo.part_vptr.as<int32_t>(), topo.part_vtx.as<int32_t>(), kpbar,
                                                 in.pvol, in.pphys, in.epsilon, in.gamma, coef);
    double* tc = static_cast<double*>(ctx->adj_tc.ensure(sizeof(double) * 12 * (size_t)M));
    k_adj_tri<<<nb256(M), 256, 0, st>>>(M, in.xyz, topo.tri.as<uint32_t>(), in.pot, in.psi, kxbar,
                                        topo.vptr.as<int32_t>(), topo.vpart.as<int32_t>(), coef, tc);
    k_adj_psibar<<<nb256(N), 256, 0, st>>>(N, in.xyz, topo.tri.as<uint32_t>(), topo.vptr.as<int32_t>(),
                                           topo.vtri.as<int32_t>(), kxbar, psibar);
    BH_HIP(hipGetLastError());
}

void adj_stage_post(bemhip_ctx* ctx, const DevTopology& topo, const AdjStageIn& in, const double* lam, double* xbar,
                    double* pbar) {
    hipStream_t st = ctx->stream;
    const int N = (int)topo.N, M = (int)topo.M;
    const int ntb = nb256(M);
    int nchunks = std::max(1, std::min((2048 + ntb - 1) / ntb, (N + ADJ_RB - 1) / ADJ_RB));
    const int rows_per_chunk = (N + nchunks - 1) / nchunks;
    nchunks = (N + rows_per_chunk - 1) / rows_per_chunk;
    double* tpart = static_cast<double*>(ctx->adj_tpart.ensure(sizeof(double) * 12 * (size_t)M * nchunks));
    double* rpart = static_cast<double*>(ctx->adj_rpart.ensure(sizeof(double) * 4 * (size_t)N * ntb));
    double* sc = static_cast<double*>(ctx->adj_sc.ensure(sizeof(double) * 27 * (size_t)M));
    k_adj_far<<<dim3(ntb, nchunks), 256, 0, st>>>(N, M, rows_per_chunk, in.xyz, topo.tri.as<uint32_t>(), in.pot,
                                                  in.psi, lam, tpart, rpart);
    k_adj_self<<<nb256(3 * (int64_t)M), 256, 0, st>>>(M, in.xyz, topo.tri.as<uint32_t>(), in.psi, lam, sc);
    k_adj_gather<<<nb256(N), 256, 0, st>>>(N, M, topo.tri.as<uint32_t>(), topo.vptr.as<int32_t>(),
                                           topo.vtri.as<int32_t>(), ctx->adj_tc.as<double>(), tpart, nchunks, rpart,
                                           ntb, sc, lam, xbar, pbar);
    BH_HIP(hipGetLastError());
}

// xbar += coef * dV/dx for the total volume (coef per part in vcoef)
void adj_volume(bemhip_ctx* ctx, const DevTopology& topo, const double* xyz, const double* vcoef, double* xbar) {
    hipStream_t st = ctx->stream;
    const int N = (int)topo.N, M = (int)topo.M;
    double* tc = static_cast<double*>(ctx->adj_tc.ensure(sizeof(double) * 12 * (size_t)M));
    double* zero = static_cast<double*>(ctx->adj_zero.ensure(sizeof(double) * 3 * (size_t)N));
    BH_HIP(hipMemsetAsync(zero, 0, sizeof(double) * 3 * (size_t)N, st));
    // with zero velocity cotangents k_adj_tri leaves only the volume term
    k_adj_tri<<<nb256(M), 256, 0, st>>>(M, xyz, topo.tri.as<uint32_t>(), zero, zero, zero, topo.vptr.as<int32_t>(),
                                        topo.vpart.as<int32_t>(), vcoef, tc);
    k_adj_gather<<<nb256(N), 256, 0, st>>>(N, M, topo.tri.as<uint32_t>(), topo.vptr.as<int32_t>(),
                                           topo.vtri.as<int32_t>(), tc, nullptr, 0, nullptr, 0, nullptr, nullptr,
                                           xbar, nullptr);
    BH_HIP(hipGetLastError());
}

}  // namespace bemhip
