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
//                column patch): vertex a, 2 ab, 2 bc, the 7 values
//                |u_q ab + v_q bc|^2 and w_q phi(y_q), unit normal, Jacobian
//                (|ab x bc|), the potential at its vertices, vertex ids and
//                patch slots -- one contiguous 256-byte record.
//   k_far        work-group = 128 collocation rows (one per lane) x a range of
//                column patches, one item of a host-built work table: one dense
//                problem for a coupled mesh, one per loose part for a batch of
//                independent bubbles (each with its own G block; about
//                ASM_WG_TARGET work-groups in all, shared in proportion to n^2).
//                A patch is PATCH (12) mesh-adjacent vertices; the
//                work-group streams the triangles incident to the patch (wave-
//                uniform, scalar loads) and accumulates each lane's 12 G (and
//                optionally H) columns in LDS, private per lane, so no atomics
//                and a fixed summation order.  The LDS per lane sets the
//                occupancy (12 columns: 6 waves per SIMD).  A triangle is an
//                entry of every patch that holds one of its vertices; each
//                entry sums only its in-patch vertices' G columns, and the
//                triangle's primary entry alone adds its whole H phi part.
//                A flushed patch column is a coalesced 1 KiB store per wave.
//                H is never stored unless asked: H phi is reduced on the fly
//                as sum_k H_k (phi_k - phi_i), which equals the 4-pi-closed
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
// the quadrature point y_q = a + u_q ab + v_q bc (Interpolator.hpp:15-21)
__host__ __device__ constexpr double q_u(int q) { return TQ_X[q] + TQ_Y[q]; }
__host__ __device__ constexpr double q_v(int q) { return TQ_Y[q]; }

// w_q phi_k(q) for a vertex index k known only at run time: a wave-uniform
// selection between the three constants (no memory access)
__device__ __forceinline__ double basis_w_sel(int q, int k) {
    return k == 0 ? basis_w(q, 0) : (k == 1 ? basis_w(q, 1) : basis_w(q, 2));
}

// The record of one patch entry (a triangle as seen from one patch), 32
// doubles, streamed with scalar loads:
//   [0..2] 2 a   [3..5] 4 ab   [6..8] 4 bc   [9..15] 4 c_q, c_q = |u_q ab + v_q bc|^2
//   [16..22] w_q phi(y_q)   [23..25] half the unit normal   [26] J = |ab x bc|
//   [27], [30], [31] phi at a, b, c   [28..29] vertex ids a, b, c and the
//   packed patch slots (bits 0-23) with the primary-entry flag (bit 24)
// With d = 2 (a - x):  4 |y_q - x|^2 = |d|^2 + 4 c_q + u_q d.(4 ab) + v_q d.(4 bc).
// The dot products are formed once per (row, entry), so a quadrature point
// costs one add and two FMAs for 4 r^2; the factors 2 and 4 (exact) make
// rsq(4 r^2) = 1 / (2 r), whose Newton step yh (3 - 4 r^2 yh^2) is 1 / r in
// three operations.
enum : int { R_A = 0, R_E1 = 3, R_E2 = 6, R_C = 9, R_WPHI = 16, R_N = 23, R_J = 26, R_PHIA = 27, R_PHIB = 30,
             R_PHIC = 31 };

// 1 / r from a4 = 4 r^2: the hardware estimate of 1 / (2 r), refined by one
// Newton step for 1 / sqrt(a4) and doubled in the same operations
__device__ __forceinline__ double inv_r_from4(double a4) {
    const double yh = __builtin_amdgcn_rsq(a4);
    return yh * fma(-(a4 * yh), yh, 3.0);
}

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
    const double pa = phi ? phi[ia] : 0.0, pb = phi ? phi[ib] : 0.0, pc = phi ? phi[ic] : 0.0;
    double* o = rec + (size_t)TQ_STRIDE * e;
    o[R_A] = 2.0 * ax; o[R_A + 1] = 2.0 * ay; o[R_A + 2] = 2.0 * az;
    o[R_E1] = 4.0 * abx; o[R_E1 + 1] = 4.0 * aby; o[R_E1 + 2] = 4.0 * abz;
    o[R_E2] = 4.0 * bcx; o[R_E2 + 1] = 4.0 * bcy; o[R_E2 + 2] = 4.0 * bcz;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double u = q_u(q), v = q_v(q);
        const double wx = u * abx + v * bcx, wy = u * aby + v * bcy, wz = u * abz + v * bcz;
        o[R_C + q] = 4.0 * (wx * wx + wy * wy + wz * wz);
        o[R_WPHI + q] = basis_w(q, 0) * pa + basis_w(q, 1) * pb + basis_w(q, 2) * pc;
    }
    o[R_N] = 0.5 * nx; o[R_N + 1] = 0.5 * ny; o[R_N + 2] = 0.5 * nz;
    o[R_J] = J;
    o[R_PHIA] = pa; o[R_PHIB] = pb; o[R_PHIC] = pc;
    uint32_t* w = reinterpret_cast<uint32_t*>(o + 28);
    w[0] = ia; w[1] = ib; w[2] = ic; w[3] = pslots[e];
}

// x's view of one entry: d = 2 (a - x) and the dot products of the expansion
// (and of the image point R x, R: x -> -x, in wall mode); x2 = 2 x
struct FarGeo {
    double D, P, Q;      // |d|^2, d.(4 ab), d.(4 bc)
    double Dm, Pm, Qm;   // the same for the image triangle R T and x
};

template <bool MIRROR>
__device__ __forceinline__ FarGeo far_geo(const double (&r)[28], double xi, double yi, double zi) {
    FarGeo f;
    const double dx = r[R_A] - xi, dy = r[R_A + 1] - yi, dz = r[R_A + 2] - zi;
    const double dyz = dy * dy + dz * dz;
    const double pyz = dy * r[R_E1 + 1] + dz * r[R_E1 + 2];
    const double qyz = dy * r[R_E2 + 1] + dz * r[R_E2 + 2];
    f.D = dx * dx + dyz;
    f.P = dx * r[R_E1] + pyz;
    f.Q = dx * r[R_E2] + qyz;
    if constexpr (MIRROR) {
        // R T has vertex R a and edges R ab, R bc: 2 (R a - x) = (-2 a_x - 2 x, ..)
        const double dxm = -r[R_A] - xi;
        f.Dm = dxm * dxm + dyz;
        f.Pm = -dxm * r[R_E1] + pyz;
        f.Qm = -dxm * r[R_E2] + qyz;
    }
    return f;
}

// One entry's 7-point sums for one collocation point x:
//   g[j] = sum_q w_q phi_k(q) / r_q for NG in-patch vertices (all three, or
//          the basis indices ka, kb);
//   SK:    s[k] = sum_q w_q phi_k(q) / r_q^3 for all three k (stored H);
//   HA:    A = sum_q w_q phi(y_q) / r_q^3 and B = sum_q w_q / r_q^3, the H phi
//          row part sum_k s_k (phi_k - phi_i) = A - phi_i B.
// The *m sums are the image point's in wall mode.
struct FarSums {
    double g[3], s[3], A, B;
    double gm[3], sm[3], Am, Bm;
};

template <bool MIRROR, int NG, bool SK, bool HA>
__device__ __forceinline__ void far_sums(const double (&r)[28], const FarGeo& f, int ka, int kb, FarSums& o) {
#pragma unroll
    for (int l = 0; l < 3; ++l) o.g[l] = o.s[l] = o.gm[l] = o.sm[l] = 0.0;
    o.A = o.B = o.Am = o.Bm = 0.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const double c = r[R_C + k];
        const double inv = inv_r_from4(fma(q_v(k), f.Q, fma(q_u(k), f.P, f.D + c)));
        if constexpr (NG == 3) {
            o.g[0] += basis_w(k, 0) * inv; o.g[1] += basis_w(k, 1) * inv; o.g[2] += basis_w(k, 2) * inv;
        } else {
            o.g[0] += basis_w_sel(k, ka) * inv;
            if constexpr (NG == 2) o.g[1] += basis_w_sel(k, kb) * inv;
        }
        if constexpr (SK || HA) {
            const double inv3 = inv * inv * inv;
            if constexpr (SK) {
                o.s[0] += basis_w(k, 0) * inv3; o.s[1] += basis_w(k, 1) * inv3; o.s[2] += basis_w(k, 2) * inv3;
            }
            if constexpr (HA) {
                o.A += r[R_WPHI + k] * inv3;
                o.B += TQ_W[k] * inv3;
            }
        }
        if constexpr (MIRROR) {
            const double invm = inv_r_from4(fma(q_v(k), f.Qm, fma(q_u(k), f.Pm, f.Dm + c)));
            if constexpr (NG == 3) {
                o.gm[0] += basis_w(k, 0) * invm; o.gm[1] += basis_w(k, 1) * invm; o.gm[2] += basis_w(k, 2) * invm;
            } else {
                o.gm[0] += basis_w_sel(k, ka) * invm;
                if constexpr (NG == 2) o.gm[1] += basis_w_sel(k, kb) * invm;
            }
            if constexpr (SK || HA) {
                const double invm3 = invm * invm * invm;
                if constexpr (SK) {
                    o.sm[0] += basis_w(k, 0) * invm3; o.sm[1] += basis_w(k, 1) * invm3;
                    o.sm[2] += basis_w(k, 2) * invm3;
                }
                if constexpr (HA) {
                    o.Am += r[R_WPHI + k] * invm3;
                    o.Bm += TQ_W[k] * invm3;
                }
            }
        }
    }
}

// (y_q - x).n for every quadrature point (they lie in the triangle's plane),
// and (y_q - R x).n for the image (image point R y with normal R n:
// (R y - x).(R n) = (y - R x).n); x2 = 2 x, the record holds 2 a and n / 2
__device__ __forceinline__ double far_zn(const double (&r)[28], double xi, double yi, double zi) {
    return (r[R_A] - xi) * r[R_N] + (r[R_A + 1] - yi) * r[R_N + 1] + (r[R_A + 2] - zi) * r[R_N + 2];
}
__device__ __forceinline__ double far_znm(const double (&r)[28], double xi, double yi, double zi) {
    return (r[R_A] + xi) * r[R_N] + (r[R_A + 1] - yi) * r[R_N + 1] + (r[R_A + 2] - zi) * r[R_N + 2];
}

// One patch entry of k_far without a stored H: G of the in-patch vertices
// into the lane's LDS columns and, on the triangle's primary entry, its whole
// H phi row contribution.  i in T: the direct part is the self panel (added
// by k_self) and H = 0; the quadrature points are strictly inside T
// (barycentric >= 0.0597), so the sums stay finite there and a zero
// Jacobian removes them.
template <bool MIRROR, bool PRIMARY, int NG>
__device__ __forceinline__ void far_entry(const double (&r)[28], const uint32_t (&v)[3], uint32_t row, double xi,
                                          double yi, double zi, double phii, const uint32_t (&sl)[3], int ka, int kb,
                                          double (*accG)[ASM_ROWS], int tid, double& hphi) {
    const FarGeo f = far_geo<MIRROR>(r, xi, yi, zi);
    FarSums o;
    far_sums<MIRROR, NG, false, PRIMARY>(r, f, ka, kb, o);
    const double J = r[R_J];
    const bool far = (row != v[0]) & (row != v[1]) & (row != v[2]);
    const double Jf = far ? J : 0.0;
    if constexpr (NG == 3) {
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double Gv = Jf * o.g[l];
            if constexpr (MIRROR) Gv += J * o.gm[l];
            accG[sl[l]][tid] += Gv;
        }
    } else {
        double Ga = Jf * o.g[0];
        if constexpr (MIRROR) Ga += J * o.gm[0];
        accG[sl[ka]][tid] += Ga;
        if constexpr (NG == 2) {
            double Gb = Jf * o.g[1];
            if constexpr (MIRROR) Gb += J * o.gm[1];
            accG[sl[kb]][tid] += Gb;
        }
    }
    if constexpr (PRIMARY) {
        // sum_k H_k (phi_k - phi_i), H_k = -((y - x).n) J s_k
        hphi += (-far_zn(r, xi, yi, zi) * Jf) * (o.A - phii * o.B);
        if constexpr (MIRROR) hphi += (-far_znm(r, xi, yi, zi) * J) * (o.Am - phii * o.Bm);
    }
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
    // twice the collocation point: the records hold 2 a (exact scalings)
    const double xi = 2.0 * xyz[3 * row], yi = 2.0 * xyz[3 * row + 1], zi = 2.0 * xyz[3 * row + 2];
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
            // the whole record before the variant is chosen: one batch of
            // scalar loads an entry
            double rv[28];
#pragma unroll
            for (int i = 0; i < 28; ++i) rv[i] = q[i];
            const uint32_t* qw = reinterpret_cast<const uint32_t*>(q + 28);
            const uint32_t vv[3] = {qw[0], qw[1], qw[2]};
            const uint32_t pk = qw[3];
            const uint32_t sl[3] = {pk & 0xFFu, (pk >> 8) & 0xFFu, (pk >> 16) & 0xFFu};
            const uint32_t r = (uint32_t)row;
            if constexpr (!STORE_H) {
                // G for the vertices in this patch only, and the triangle's
                // H phi once, on its primary entry
                const bool primary = (pk >> 24) & 1u;
                const int ng = (sl[0] != 0xFFu) + (sl[1] != 0xFFu) + (sl[2] != 0xFFu);
                // the in-patch basis indices when not all three are
                const int ka = sl[0] != 0xFFu ? 0 : (sl[1] != 0xFFu ? 1 : 2);
                const int kb = ng == 2 ? (sl[2] != 0xFFu ? 2 : 1) : ka;
                if (primary) {
                    if (ng == 3) far_entry<MIRROR, true, 3>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                    else if (ng == 2) far_entry<MIRROR, true, 2>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                    else far_entry<MIRROR, true, 1>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                } else {
                    if (ng == 3) far_entry<MIRROR, false, 3>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                    else if (ng == 2) far_entry<MIRROR, false, 2>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                    else far_entry<MIRROR, false, 1>(rv, vv, r, xi, yi, zi, phii, sl, ka, kb, accG, tid, hphi);
                }
            } else {
                // stored H: all three G and H of the entry, each in-patch
                // vertex's into its columns, H phi and the H row sum per vertex
                const FarGeo f = far_geo<MIRROR>(rv, xi, yi, zi);
                FarSums o;
                far_sums<MIRROR, 3, true, false>(rv, f, 0, 0, o);
                const bool far = (r != vv[0]) & (r != vv[1]) & (r != vv[2]);
                const double J = rv[R_J];
                const double Jf = far ? J : 0.0;
                const double Hf = -far_zn(rv, xi, yi, zi) * Jf;
                const double Hm = MIRROR ? -far_znm(rv, xi, yi, zi) * J : 0.0;
                const double phk[3] = {rv[R_PHIA], q[R_PHIB], q[R_PHIC]};
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    if (sl[l] == 0xFFu) continue;
                    double Gv = Jf * o.g[l], Hv = Hf * o.s[l];
                    if constexpr (MIRROR) {
                        Gv += J * o.gm[l];
                        Hv += Hm * o.sm[l];
                    }
                    accG[sl[l]][tid] += Gv;
                    accH[sl[l]][tid] += Hv;
                    hphi += Hv * (phk[l] - phii);
                    hrow += Hv;
                }
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
