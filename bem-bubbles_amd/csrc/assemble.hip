// Collocation assembly of the single-layer (G) and double-layer (H) matrices
// for linear elements on flat triangles, gfx950.
//
// Reference semantics: ColocSim::assemble_matrices (ColocSim.cpp:21-110) with
// integrate_Lin_coloc_local[_mir] (Integrator.cpp:237-283): for every
// collocation vertex i and triangle T, 3 G and 3 H panel integrals are added to
// the columns of T's vertices; i in T uses the polar self-panel rule for G and
// H = 0 (plus the image panel in wall mode); then every H row is closed by the
// 4-pi rule H(i,i) = -4 pi - sum_{j != i} H(i,j).
//
// Device design:
//   k_entry_prep one thread per patch entry (a triangle as seen from one
//                column patch): the 7 quadrature points, unit normal, Jacobian
//                (|ab x bc|), the potential at its vertices, vertex ids and
//                patch slots -- one contiguous 256-byte record.
//   k_far        work-group = 128 collocation rows (one per lane) x a range of
//                column patches, one item of a host-built work table: one dense
//                problem for a coupled mesh, one per loose part for a batch of
//                independent bubbles (each with its own G block, about 8192
//                work-groups in all, shared in proportion to n^2).
//                A patch is PATCH (16) mesh-adjacent vertices; the
//                work-group streams the triangles incident to the patch (wave-
//                uniform, scalar loads) and accumulates each lane's 16 G (and
//                optionally H) columns in LDS, private per lane, so no atomics
//                and a fixed summation order.  The LDS per lane sets the
//                occupancy: 16 columns allow 5 waves per SIMD, 32 only 2.5
//                (measured 25% slower despite 11% fewer triangle entries).  A flushed patch column is a
//                coalesced 1 KiB store per wave.  H is never stored unless
//                asked: H phi is reduced on the fly as
//                sum_k H_k (phi_k - phi_i), which equals the 4-pi-closed
//                product (H phi)_i + 4 pi phi_i exactly in real arithmetic.
//   k_self       one thread per vertex: adds the polar-rule self panels to its
//                own row, finishes H phi (and H's diagonal when stored).
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "quad_rules.hpp"
#include "vec3.hpp"

namespace bemhip {

constexpr int TQ_STRIDE = 32;  // doubles per triangle record

// w_q * phi_k(q), phi = (1 - x - y, x, y): basis of get_linear_elements
// (ResultTypes.hpp:111-117) at u = x + y, v = y.
__host__ __device__ constexpr double basis_w(int q, int k) {
    return TQ_W[q] * (k == 0 ? (1.0 - (TQ_X[q] + TQ_Y[q])) : (k == 1 ? TQ_X[q] : TQ_Y[q]));
}

// One record per patch entry (a triangle as seen from one patch), so the
// assembly streams records contiguously with a single scalar load per
// triangle: the 7 quadrature points, unit normal, Jacobian, the potential at
// the 3 vertices, the vertex ids and the vertex slots in this patch.
__global__ __launch_bounds__(256) void k_entry_prep(int nent, const double* __restrict__ xyz,
                                                    const double* __restrict__ phi, const uint32_t* __restrict__ tri,
                                                    const uint32_t* __restrict__ ptri,
                                                    const uint32_t* __restrict__ pslots, double* __restrict__ rec) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const uint32_t j = ptri[e];
    const uint32_t ia = tri[3 * j], ib = tri[3 * j + 1], ic = tri[3 * j + 2];
    const double ax = xyz[3 * ia], ay = xyz[3 * ia + 1], az = xyz[3 * ia + 2];
    const double bx = xyz[3 * ib], by = xyz[3 * ib + 1], bz = xyz[3 * ib + 2];
    const double cx = xyz[3 * ic], cy = xyz[3 * ic + 1], cz = xyz[3 * ic + 2];
    // Interpolator (Interpolator.hpp:15-21): ab, bc, n = ab x bc, area_ = |n|
    const double abx = bx - ax, aby = by - ay, abz = bz - az;
    const double bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
    double nx = aby * bcz - abz * bcy, ny = abz * bcx - abx * bcz, nz = abx * bcy - aby * bcx;
    const double J = sqrt(nx * nx + ny * ny + nz * nz);
    const double rj = 1.0 / J;
    nx *= rj; ny *= rj; nz *= rj;
    double* o = rec + (size_t)TQ_STRIDE * e;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double u = TQ_X[q] + TQ_Y[q], v = TQ_Y[q];
        o[3 * q + 0] = ax + (u * abx + v * bcx);
        o[3 * q + 1] = ay + (u * aby + v * bcy);
        o[3 * q + 2] = az + (u * abz + v * bcz);
    }
    o[21] = nx; o[22] = ny; o[23] = nz;
    o[24] = J;
    o[25] = phi ? phi[ia] : 0.0;
    o[26] = phi ? phi[ib] : 0.0;
    o[27] = phi ? phi[ic] : 0.0;
    uint32_t* w = reinterpret_cast<uint32_t*>(o + 28);
    w[0] = ia; w[1] = ib; w[2] = ic; w[3] = pslots[e];
    o[30] = 0.0; o[31] = 0.0;
}

template <bool MIRROR, bool STORE_H, bool WRITE32>
__global__ __launch_bounds__(ASM_ROWS) void k_far(
    int Ntot, const AsmItem* __restrict__ items, const AsmProb* __restrict__ probs, const double* __restrict__ xyz,
    const double* __restrict__ phi, const double* __restrict__ rec, const int32_t* __restrict__ pptr,
    const int32_t* __restrict__ pcol, double* __restrict__ G, float* __restrict__ G32, double* __restrict__ H,
    double* __restrict__ part_hphi, double* __restrict__ part_hrow) {
    __shared__ double accG[PATCH][ASM_ROWS];
    __shared__ double accH[STORE_H ? PATCH : 1][ASM_ROWS];
    const AsmItem it = items[blockIdx.x];
    const AsmProb pr = probs[it.prob];
    const int tid = threadIdx.x;
    const int rloc = it.row0 + tid;                 // row inside the problem
    const bool valid = rloc < pr.n;
    const int row = pr.v0 + (valid ? rloc : pr.n - 1);   // global vertex id
    const size_t ld = (size_t)pr.n;
    const double xi = xyz[3 * row], yi = xyz[3 * row + 1], zi = xyz[3 * row + 2];
    const double phii = phi ? phi[row] : 0.0;
    double hphi = 0.0, hrow = 0.0;

    const int p0 = it.p0, p1 = it.p1;
    for (int p = p0; p < p1; ++p) {
#pragma unroll
        for (int s = 0; s < PATCH; ++s) {
            accG[s][tid] = 0.0;
            if constexpr (STORE_H) accH[s][tid] = 0.0;
        }
        const int e0 = pptr[p], e1 = pptr[p + 1];
        for (int e = e0; e < e1; ++e) {
            const double* __restrict__ q = rec + (size_t)TQ_STRIDE * e;
            const uint32_t* w = reinterpret_cast<const uint32_t*>(q + 28);
            const uint32_t v0 = w[0], v1 = w[1], v2 = w[2], sl = w[3];
            const double nx = q[21], ny = q[22], nz = q[23], J = q[24];
            // every quadrature point lies in the triangle's plane: (y_q - x).n is one number
            const double dy = q[1] - yi, dz = q[2] - zi;
            const double zn = (q[0] - xi) * nx + dy * ny + dz * nz;
            double g0 = 0, g1 = 0, g2 = 0, s0 = 0, s1 = 0, s2 = 0;
            double gm0 = 0, gm1 = 0, gm2 = 0, sm0 = 0, sm1 = 0, sm2 = 0;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const double zx = q[3 * k] - xi, zy = q[3 * k + 1] - yi, zz = q[3 * k + 2] - zi;
                const double yz2 = zy * zy + zz * zz;
                const double inv = rsqrt_nr(zx * zx + yz2);
                const double inv3 = inv * inv * inv;
                g0 += basis_w(k, 0) * inv; g1 += basis_w(k, 1) * inv; g2 += basis_w(k, 2) * inv;
                s0 += basis_w(k, 0) * inv3; s1 += basis_w(k, 1) * inv3; s2 += basis_w(k, 2) * inv3;
                if constexpr (MIRROR) {
                    const double zxm = -q[3 * k] - xi;
                    const double invm = rsqrt_nr(zxm * zxm + yz2);
                    const double invm3 = invm * invm * invm;
                    gm0 += basis_w(k, 0) * invm; gm1 += basis_w(k, 1) * invm; gm2 += basis_w(k, 2) * invm;
                    sm0 += basis_w(k, 0) * invm3; sm1 += basis_w(k, 1) * invm3; sm2 += basis_w(k, 2) * invm3;
                }
            }
            // i in T: the direct part is the self panel (added by k_self) and H = 0.
            // The quadrature points are strictly inside T (barycentric >= 0.0597),
            // so g, s stay finite there and a zero Jacobian removes them.
            const bool far = ((uint32_t)row != v0) & ((uint32_t)row != v1) & ((uint32_t)row != v2);
            const double Jf = far ? J : 0.0;
            const double Hf = -zn * Jf;
            double G0 = Jf * g0, G1 = Jf * g1, G2 = Jf * g2;
            double H0 = Hf * s0, H1 = Hf * s1, H2 = Hf * s2;
            if constexpr (MIRROR) {
                // image point R y (R: x -> -x) with normal R n: (Ry - x).(Rn) = (y - Rx).n
                const double znm = (q[0] + xi) * nx + dy * ny + dz * nz;
                const double Hm = -znm * J;
                G0 += J * gm0; G1 += J * gm1; G2 += J * gm2;
                H0 += Hm * sm0; H1 += Hm * sm1; H2 += Hm * sm2;
            }
            const uint32_t sa = sl & 0xFFu, sb = (sl >> 8) & 0xFFu, sc = (sl >> 16) & 0xFFu;
            if (sa != 0xFFu) {
                accG[sa][tid] += G0;
                if constexpr (STORE_H) accH[sa][tid] += H0;
                hphi += H0 * (q[25] - phii);
                if constexpr (STORE_H) hrow += H0;
            }
            if (sb != 0xFFu) {
                accG[sb][tid] += G1;
                if constexpr (STORE_H) accH[sb][tid] += H1;
                hphi += H1 * (q[26] - phii);
                if constexpr (STORE_H) hrow += H1;
            }
            if (sc != 0xFFu) {
                accG[sc][tid] += G2;
                if constexpr (STORE_H) accH[sc][tid] += H2;
                hphi += H2 * (q[27] - phii);
                if constexpr (STORE_H) hrow += H2;
            }
        }
        if (valid) {
            const int32_t* cols = pcol + (size_t)p * PATCH;
#pragma unroll 4
            for (int s = 0; s < PATCH; ++s) {
                const int32_t c = cols[s];
                if (c < 0) continue;
                const size_t off = (size_t)pr.gofs + (size_t)(c - pr.v0) * ld + (size_t)rloc;
                const double gv = accG[s][tid];
                if (G) G[off] = gv;
                if constexpr (WRITE32) G32[off] = (float)gv;
                if constexpr (STORE_H) H[off] = accH[s][tid];
            }
        }
    }
    if (valid) {
        part_hphi[(size_t)it.g * Ntot + row] = hphi;
        if constexpr (STORE_H) part_hrow[(size_t)it.g * Ntot + row] = hrow;
    }
}

// integrate_identical_coloc<LinElm> (Integrator.cpp:12-41): x = a, polar rule
// theta in (0, pi/4) with the 7-point Gauss rule.
__device__ inline void self_panel(double abx, double aby, double abz, double bcx, double bcy, double bcz,
                                  double J, double& t0, double& t1, double& t2) {
    const double jac = 0.5 * J * (PI / 4.0);
    t0 = t1 = t2 = 0.0;
#pragma unroll
    for (int p = 0; p < 7; ++p) {
        const double th = GQ_X[p] * (PI / 4.0);
        const double c = cos(th), s = sin(th);
        const double dx = c * abx + s * bcx, dy = c * aby + s * bcy, dz = c * abz + s * bcz;
        const double f = jac * GQ_W[p] * rsqrt(dx * dx + dy * dy + dz * dz);
        const double fac = 1.0 / (c * c);
        t0 += f * c * fac;
        t1 += f * (c - s) * fac;
        t2 += f * s * fac;
    }
}

template <bool STORE_H, bool WRITE32>
__global__ __launch_bounds__(256) void k_self(int N, const int32_t* __restrict__ vprob,
                                              const AsmProb* __restrict__ probs, const double* __restrict__ xyz,
                                              const double* __restrict__ phi, const uint32_t* __restrict__ tri,
                                              const int32_t* __restrict__ vptr, const int32_t* __restrict__ vtri,
                                              double* __restrict__ G, float* __restrict__ G32,
                                              double* __restrict__ H, const double* __restrict__ part_hphi,
                                              const double* __restrict__ part_hrow, double* __restrict__ hphi) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const AsmProb pr = probs[vprob ? vprob[v] : 0];
    const size_t ld = (size_t)pr.n;
    const size_t rloc = (size_t)(v - pr.v0);
    const int ngroups = pr.groups;
    const double ax = xyz[3 * v], ay = xyz[3 * v + 1], az = xyz[3 * v + 2];
    for (int e = vptr[v]; e < vptr[v + 1]; ++e) {
        const uint32_t* t = tri + 3 * (size_t)vtri[e];
        // cyclic_reorder (Triplet.hpp:55-72): v first
        const int s = (t[0] == (uint32_t)v) ? 0 : ((t[1] == (uint32_t)v) ? 1 : 2);
        const uint32_t ib = t[(s + 1) % 3], ic = t[(s + 2) % 3];
        const double bx = xyz[3 * ib], by = xyz[3 * ib + 1], bz = xyz[3 * ib + 2];
        const double cx = xyz[3 * ic], cy = xyz[3 * ic + 1], cz = xyz[3 * ic + 2];
        const double abx = bx - ax, aby = by - ay, abz = bz - az;
        const double bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
        const double nx = aby * bcz - abz * bcy, ny = abz * bcx - abx * bcz, nz = abx * bcy - aby * bcx;
        const double J = sqrt(nx * nx + ny * ny + nz * nz);
        double g0, g1, g2;
        self_panel(abx, aby, abz, bcx, bcy, bcz, J, g0, g1, g2);
        const uint32_t cols[3] = {(uint32_t)v, ib, ic};
        const double gv[3] = {g0, g1, g2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t off = (size_t)pr.gofs + (size_t)(cols[k] - (uint32_t)pr.v0) * ld + rloc;
            double val = gv[k];
            if (G) { val += G[off]; G[off] = val; }
            else val += (double)G32[off];
            if constexpr (WRITE32) G32[off] = (float)val;
        }
    }
    double hp = 0.0, hr = 0.0;
    for (int g = 0; g < ngroups; ++g) {
        hp += part_hphi[(size_t)g * N + v];
        if constexpr (STORE_H) hr += part_hrow[(size_t)g * N + v];
    }
    const double four_pi = 4.0 * PI;
    if (hphi) hphi[v] = hp - four_pi * (phi ? phi[v] : 0.0);
    if constexpr (STORE_H) {
        const size_t d = (size_t)pr.gofs + rloc * ld + rloc;
        H[d] = H[d] - four_pi - hr;  // ColocSim.cpp:96 with val_H = -sum_j H(i,j)
    }
}

void assemble_device(bemhip_ctx* ctx, const DevTopology& topo, const double* d_xyz, const double* d_phi,
                     uint32_t flags, const AsmOut& out) {
    const int N = (int)topo.N;
    BH_CHECK(out.G64 || out.G32, BEMHIP_ERR_ARG, "assembly needs a G output");
    hipStream_t st = ctx->stream;
    const int nent = topo.nentries;
    double* rec = static_cast<double*>(ctx->triq.ensure(sizeof(double) * TQ_STRIDE * (size_t)nent));
    k_entry_prep<<<(nent + 255) / 256, 256, 0, st>>>(nent, d_xyz, d_phi, topo.tri.as<uint32_t>(),
                                                      topo.patch_tri.as<uint32_t>(), topo.patch_slots.as<uint32_t>(),
                                                      rec);
    BH_HIP(hipGetLastError());

    const int groups = topo.max_groups;
    double* ph = static_cast<double*>(ctx->part_hphi.ensure(sizeof(double) * (size_t)groups * N));
    double* hr = static_cast<double*>(ctx->part_hrow.ensure(sizeof(double) * (size_t)groups * N));
    const bool mirror = flags & BEMHIP_MIRROR_X0;
    const bool storeH = out.H != nullptr;
    const bool w32 = out.G32 != nullptr;
    const AsmItem* items = topo.d_items.as<AsmItem>();
    const AsmProb* probs = topo.d_probs.as<AsmProb>();
    dim3 grid(topo.nitems);
#define BH_FAR(MI, SH, W3)                                                                                 \
    k_far<MI, SH, W3><<<grid, ASM_ROWS, 0, st>>>(N, items, probs, d_xyz, d_phi, rec,                      \
                                                 topo.patch_ptr.as<int32_t>(), topo.patch_col.as<int32_t>(), \
                                                 out.G64, out.G32, out.H, ph, hr)
    if (mirror) {
        if (storeH) { if (w32) BH_FAR(true, true, true); else BH_FAR(true, true, false); }
        else { if (w32) BH_FAR(true, false, true); else BH_FAR(true, false, false); }
    } else {
        if (storeH) { if (w32) BH_FAR(false, true, true); else BH_FAR(false, true, false); }
        else { if (w32) BH_FAR(false, false, true); else BH_FAR(false, false, false); }
    }
#undef BH_FAR
    BH_HIP(hipGetLastError());
#define BH_SELF(SH, W3)                                                                             \
    k_self<SH, W3><<<(N + 255) / 256, 256, 0, st>>>(N, topo.independent ? topo.vpart.as<int32_t>() : nullptr, \
                                                    probs, d_xyz, d_phi, topo.tri.as<uint32_t>(),                  \
                                                    topo.vptr.as<int32_t>(), topo.vtri.as<int32_t>(), \
                                                    out.G64, out.G32, out.H, ph, hr, out.hphi)
    if (storeH) { if (w32) BH_SELF(true, true); else BH_SELF(true, false); }
    else { if (w32) BH_SELF(false, true); else BH_SELF(false, false); }
#undef BH_SELF
    BH_HIP(hipGetLastError());
}

}  // namespace bemhip
