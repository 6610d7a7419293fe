// Potential at points off the surface from the boundary data (phi, psi):
// compute_exterior_pot (LinLinSim.cpp:440-477) over
// Integrator::get_exterior_potential (Integrator.cpp:287-305):
//
//   phi_ext(y) = 1/(4 pi) sum_T [ -sum_k G_k(y, T) psi_k + sum_k H_k(y, T) phi_k ]
//
// with G_k, H_k the disjoint-panel integrals of integrate_disjoint_coloc
// (Integrator.hpp:471-491) under the 19-point degree-9 rule (no mirror image).
// Per triangle the basis sums fold into the quadrature weights,
//   sum_k G_k psi_k = J sum_q w_q psi(q) / r_q,
//   sum_k H_k phi_k = J sum_q w_q phi(q) (-(y_q - y).n) / r_q^3,
// and (y_q - y).n is one number per triangle (every point lies in its plane).
//
// Device design:
//   k_ext_prep  one thread per triangle: its 19 points and the weights
//               J w_q psi(q), J w_q phi(q), its normal and first corner -- one
//               contiguous record.
//   k_ext       a work-group of 256 targets x a chunk of triangles; records
//               are staged through LDS 8 at a time; each chunk writes its
//               partial sums, which k_ext_sum adds in a fixed order
//               (deterministic).
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "quad_rules.hpp"
#include "vec3.hpp"

namespace bemhip {

constexpr int EXT_Q = 19;
constexpr int EXT_REC = 104;      // doubles per record: 19 x (x, y, z, wpsi, wphi), normal, corner
constexpr int EXT_TGT = 256;      // targets per work-group

__global__ __launch_bounds__(256) void k_ext_prep(int64_t M, const double* __restrict__ xyz,
                                                  const uint32_t* __restrict__ tri, const double* __restrict__ phi,
                                                  const double* __restrict__ psi, double* __restrict__ rec) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    const uint32_t ia = tri[3 * t], ib = tri[3 * t + 1], ic = tri[3 * t + 2];
    const d3 a = ld3(xyz + 3 * (size_t)ia), b = ld3(xyz + 3 * (size_t)ib), c = ld3(xyz + 3 * (size_t)ic);
    // Interpolator (Interpolator.hpp:15-21): ab, bc, n = ab x bc / |ab x bc|, J = |ab x bc|
    const d3 ab = sub(b, a), bc = sub(c, b);
    d3 n = crs(ab, bc);
    const double J = sqrt(dot(n, n));
    n = scl(1.0 / J, n);
    double* o = rec + (size_t)EXT_REC * t;
#pragma unroll
    for (int q = 0; q < EXT_Q; ++q) {
        // the reference's quadrature point (x, y) on the other unit triangle:
        // u = x + y, v = y; basis [1 - u, u - v, v] (ResultTypes.hpp:111-117)
        const double u = Q19_X[q] + Q19_Y[q], v = Q19_Y[q];
        const d3 y = add(a, add(scl(u, ab), scl(v, bc)));
        const double f0 = 1.0 - u, f1 = u - v, f2 = v;
        const double wj = J * Q19_W[q];
        o[5 * q + 0] = y.x;
        o[5 * q + 1] = y.y;
        o[5 * q + 2] = y.z;
        o[5 * q + 3] = wj * (f0 * psi[ia] + f1 * psi[ib] + f2 * psi[ic]);
        o[5 * q + 4] = wj * (f0 * phi[ia] + f1 * phi[ib] + f2 * phi[ic]);
    }
    o[95] = n.x; o[96] = n.y; o[97] = n.z;
    o[98] = a.x; o[99] = a.y; o[100] = a.z;
}

// part[chunk * P + i] = sum over the chunk's triangles for target i
// Every lane of the work-group walks the same triangles: the record address
// is wave-uniform, so its values come in through scalar loads and feed the
// FP64 arithmetic as scalar operands (no LDS staging, no barriers).  65536
// targets x 81920 panels: 1.7e12 kernel evaluations per second, 24 TFLOP/s.
__global__ __launch_bounds__(EXT_TGT) void k_ext(int64_t P, int64_t M, int tpc, const double* __restrict__ tgt,
                                                 const double* __restrict__ rec, double* __restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * EXT_TGT + threadIdx.x;
    const int64_t ic = i < P ? i : P - 1;
    const double yx = tgt[3 * ic], yy = tgt[3 * ic + 1], yz = tgt[3 * ic + 2];
    const int64_t t0 = (int64_t)blockIdx.y * tpc, t1 = t0 + tpc < M ? t0 + tpc : M;
    double acc = 0.0;
    {
        for (int64_t k = t0; k < t1; ++k) {
            const double* __restrict__ r = rec + (size_t)EXT_REC * k;
            // (y_q - y).n, the same at every point of the plane
            const double zn = (r[98] - yx) * r[95] + (r[99] - yy) * r[96] + (r[100] - yz) * r[97];
            double gs = 0.0, hs = 0.0;
#pragma unroll
            for (int q = 0; q < EXT_Q; ++q) {
                const double dx = r[5 * q] - yx, dy = r[5 * q + 1] - yy, dz = r[5 * q + 2] - yz;
                const double inv = rsqrt_nr(dx * dx + dy * dy + dz * dz);
                gs += r[5 * q + 3] * inv;
                hs += r[5 * q + 4] * (inv * inv * inv);
            }
            // -sum_k G_k psi_k + sum_k H_k phi_k, H's integrand being -(y_q - y).n / r^3
            acc += -gs - zn * hs;
        }
    }
    if (i < P) part[(size_t)blockIdx.y * P + i] = acc;
}

__global__ void k_ext_sum(int64_t P, int chunks, const double* __restrict__ part, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * P + i];
    out[i] = s * (1.0 / (4.0 * PI));
}

void exterior_device(bemhip_ctx* ctx, int64_t M, const double* xyz, const uint32_t* tri, const double* phi,
                     const double* psi, int64_t P, const double* tgt, double* out) {
    if (P <= 0) return;
    hipStream_t st = ctx->stream;
    double* rec = static_cast<double*>(ctx->ext_rec.ensure(sizeof(double) * EXT_REC * (size_t)std::max<int64_t>(M, 1)));
    if (M > 0) {
        k_ext_prep<<<(unsigned)((M + 255) / 256), 256, 0, st>>>(M, xyz, tri, phi, psi, rec);
        BH_HIP(hipGetLastError());
    }
    // enough work-groups to fill the chip: split the triangles into chunks
    const int64_t gx = (P + EXT_TGT - 1) / EXT_TGT;
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>((8 * ctx->cus + gx - 1) / gx, (M + 63) / 64));
    chunks = std::min<int64_t>(chunks, 65535);
    const int tpc = (int)((M + chunks - 1) / std::max<int64_t>(chunks, 1));
    chunks = M > 0 ? (M + tpc - 1) / tpc : 1;
    double* part = static_cast<double*>(ctx->ext_part.ensure(sizeof(double) * (size_t)chunks * P));
    if (M > 0) {
        k_ext<<<dim3((unsigned)gx, (unsigned)chunks), EXT_TGT, 0, st>>>(P, M, tpc, tgt, rec, part);
        BH_HIP(hipGetLastError());
    } else {
        BH_HIP(hipMemsetAsync(part, 0, sizeof(double) * P, st));
    }
    k_ext_sum<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(P, (int)chunks, part, out);
    BH_HIP(hipGetLastError());
}

// ------------------------------------------------ no-allocation variant --
// The work buffers of exterior_device (triangle records, per-chunk partial
// sums) and an error word in one fixed device array per GPU, for callers that
// cannot allocate: records first (EXT_REC doubles per triangle), partial sums
// after them, the error word in the last slot.  The array's host-side name is
// not a device address: exterior_scratch_init takes its address on the
// current device once (at context creation) and every use goes through that
// pointer.  Calls that overlap on one GPU share the array and race.
constexpr size_t EXT_SCRATCH = (size_t)8 << 20;    // doubles (64 MiB)
constexpr size_t EXT_SCRATCH_USABLE = EXT_SCRATCH - 32;
static __device__ __align__(256) double scratch[EXT_SCRATCH];

void exterior_scratch_init(bemhip_ctx* ctx) {
    void* p = nullptr;
    BH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BH_CHECK(p != nullptr, BEMHIP_ERR_HIP, "no device address for the exterior-potential scratch array");
    ctx->ext_scratch = static_cast<double*>(p);
}

// flag |= 1: a triangle index >= N; |= 2: a non-finite phi, psi or point
__global__ __launch_bounds__(256) void k_ext_check(int64_t N, int64_t M, const uint32_t* __restrict__ tri,
                                                   const double* __restrict__ phi, const double* __restrict__ psi,
                                                   int64_t P, const double* __restrict__ pts,
                                                   unsigned int* __restrict__ flag) {
    unsigned int f = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * M; i += stride)
        if (tri[i] >= (uint64_t)N) f |= 1u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride)
        if (!isfinite(phi[i]) || !isfinite(psi[i])) f |= 2u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * P; i += stride)
        if (!isfinite(pts[i])) f |= 2u;
    if (f) atomicOr(flag, f);
}

__global__ void k_ext_flag_clear(unsigned int* flag) {
    if (threadIdx.x == 0) *flag = 0u;
}

void exterior_device_noalloc(bemhip_ctx* ctx, int64_t N, int64_t M, const double* xyz, const uint32_t* tri,
                             const double* phi, const double* psi, int64_t P, const double* tgt, double* out) {
    BH_CHECK(ctx->ext_scratch != nullptr, BEMHIP_ERR_HIP, "exterior-potential scratch array not set up");
    BH_CHECK((size_t)EXT_REC * M + (size_t)P <= EXT_SCRATCH_USABLE, BEMHIP_ERR_ARG,
             "104 M + P exceeds the no-allocation scratch (8388576 doubles)");
    hipStream_t st = ctx->stream;
    double* rec = ctx->ext_scratch;
    unsigned int* flag = reinterpret_cast<unsigned int*>(ctx->ext_scratch + EXT_SCRATCH - 1);
    k_ext_flag_clear<<<1, 64, 0, st>>>(flag);
    k_ext_check<<<256, 256, 0, st>>>(N, M, tri, phi, psi, P, tgt, flag);
    BH_HIP(hipGetLastError());
    unsigned int* hflag = reinterpret_cast<unsigned int*>(ctx->host_scal + 48);
    BH_HIP(hipMemcpyAsync(hflag, flag, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    BH_HIP(hipStreamSynchronize(st));
    BH_CHECK(!(*hflag & 1u), BEMHIP_ERR_BAD_MESH, "vertex index out of range");
    BH_CHECK(!(*hflag & 2u), BEMHIP_ERR_NAN, "non-finite phi, psi or point");
    k_ext_prep<<<(unsigned)((M + 255) / 256), 256, 0, st>>>(M, xyz, tri, phi, psi, rec);
    BH_HIP(hipGetLastError());
    // as many triangle chunks as fill the chip and fit beside the records
    const int64_t gx = (P + EXT_TGT - 1) / EXT_TGT;
    const int64_t room = (int64_t)((EXT_SCRATCH_USABLE - (size_t)EXT_REC * M) / (size_t)P);
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>((8 * ctx->cus + gx - 1) / gx, (M + 63) / 64));
    chunks = std::min<int64_t>(std::min<int64_t>(chunks, room), 65535);
    const int tpc = (int)((M + chunks - 1) / chunks);
    chunks = (M + tpc - 1) / tpc;
    double* part = rec + (size_t)EXT_REC * M;
    k_ext<<<dim3((unsigned)gx, (unsigned)chunks), EXT_TGT, 0, st>>>(P, M, tpc, tgt, rec, part);
    k_ext_sum<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(P, (int)chunks, part, out);
    BH_HIP(hipGetLastError());
    BH_HIP(hipStreamSynchronize(st));
}

}  // namespace bemhip
