// Host-side topology tables of a mesh: assembly patches, vertex incidence and
// loose parts.  They depend on connectivity only (positions merely seed the
// patch ordering), so they are built once per mesh and reused every stage.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.hpp"

namespace bemhip {

static uint32_t spread10(uint32_t v) {
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

Topology build_topology(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, bool independent) {
    BH_CHECK(N > 0 && M > 0, BEMHIP_ERR_ARG, "empty mesh");
    BH_CHECK(N < (int64_t)1 << 31 && 3 * M < (int64_t)1 << 31, BEMHIP_ERR_ARG, "mesh too large");
    Topology T;
    T.N = N;
    T.M = M;
    T.tri.assign(tri, tri + 3 * M);
    for (int64_t j = 0; j < M; ++j) {
        const uint32_t a = tri[3 * j], b = tri[3 * j + 1], c = tri[3 * j + 2];
        BH_CHECK(a < N && b < N && c < N, BEMHIP_ERR_BAD_MESH,
                 "triangle " + std::to_string(j) + " has a vertex index out of range");
        BH_CHECK(a != b && b != c && a != c, BEMHIP_ERR_BAD_MESH,
                 "triangle " + std::to_string(j) + " repeats a vertex");
    }
    for (int64_t i = 0; i < 3 * N; ++i)
        BH_CHECK(std::isfinite(xyz[i]), BEMHIP_ERR_NAN, "non-finite vertex coordinate");

    // vertex -> triangles, ascending (Mesh.cpp:75-84)
    T.vptr.assign(N + 1, 0);
    for (int64_t j = 0; j < 3 * M; ++j) T.vptr[tri[j] + 1]++;
    for (int64_t i = 0; i < N; ++i) T.vptr[i + 1] += T.vptr[i];
    T.vtri.resize(3 * M);
    {
        std::vector<int32_t> fill(T.vptr.begin(), T.vptr.end() - 1);
        for (int64_t j = 0; j < M; ++j)
            for (int k = 0; k < 3; ++k) T.vtri[fill[tri[3 * j + k]]++] = (int32_t)j;
    }
    for (int64_t i = 0; i < N; ++i)
        BH_CHECK(T.vptr[i + 1] > T.vptr[i], BEMHIP_ERR_BAD_MESH,
                 "vertex " + std::to_string(i) + " belongs to no triangle");

    // loose parts by union-find; part ids ordered by their smallest vertex
    {
        std::vector<int32_t> p(N);
        std::iota(p.begin(), p.end(), 0);
        auto find = [&](int32_t a) {
            while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
            return a;
        };
        for (int64_t j = 0; j < M; ++j)
            for (int k = 1; k < 3; ++k) {
                int32_t a = find(tri[3 * j]), b = find(tri[3 * j + k]);
                if (a != b) { if (a < b) p[b] = a; else p[a] = b; }
            }
        std::vector<int32_t> label(N, -1);
        T.vpart.resize(N);
        int n = 0;
        for (int64_t i = 0; i < N; ++i) {
            int32_t r = find((int32_t)i);
            if (label[r] < 0) label[r] = n++;
            T.vpart[i] = label[r];
        }
        T.nparts = n;
        T.part_tptr.assign(n + 1, 0);
        for (int64_t j = 0; j < M; ++j) T.part_tptr[T.vpart[tri[3 * j]] + 1]++;
        for (int c = 0; c < n; ++c) T.part_tptr[c + 1] += T.part_tptr[c];
        T.part_tri.resize(M);
        std::vector<int32_t> fill(T.part_tptr.begin(), T.part_tptr.end() - 1);
        for (int64_t j = 0; j < M; ++j) T.part_tri[fill[T.vpart[tri[3 * j]]]++] = (int32_t)j;
        T.part_vptr.assign(n + 1, 0);
        for (int64_t i = 0; i < N; ++i) T.part_vptr[T.vpart[i] + 1]++;
        for (int c = 0; c < n; ++c) T.part_vptr[c + 1] += T.part_vptr[c];
        T.part_vtx.resize(N);
        std::vector<int32_t> vfill(T.part_vptr.begin(), T.part_vptr.end() - 1);
        for (int64_t i = 0; i < N; ++i) T.part_vtx[vfill[T.vpart[i]]++] = (int32_t)i;
    }

    // assembly patches: compact groups of PATCH vertices grown breadth-first
    // from seeds taken in Morton order, so a patch touches few triangles that
    // also feed another patch.
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], xyz[3 * i + d]);
                hi[d] = std::max(hi[d], xyz[3 * i + d]);
            }
        std::vector<std::pair<uint32_t, int32_t>> code(N);
        for (int64_t i = 0; i < N; ++i) {
            uint32_t q[3];
            for (int d = 0; d < 3; ++d) {
                const double span = hi[d] - lo[d];
                double u = span > 0 ? (xyz[3 * i + d] - lo[d]) / span : 0.0;
                q[d] = (uint32_t)std::min(1023.0, std::max(0.0, u * 1023.0));
            }
            code[i] = {spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2), (int32_t)i};
        }
        std::sort(code.begin(), code.end());
        std::vector<int32_t> pid(N, -1), slot(N, -1);
        std::vector<int32_t> queue;
        queue.reserve(N);
        int np = 0;
        for (int64_t s = 0; s < N; ++s) {
            const int32_t seed = code[s].second;
            if (pid[seed] >= 0) continue;
            int filled = 0;
            queue.clear();
            queue.push_back(seed);
            pid[seed] = np;
            size_t head = 0;
            while (head < queue.size() && filled < PATCH) {
                const int32_t v = queue[head++];
                slot[v] = filled++;
                T.patch_col.push_back(v);
                if (filled == PATCH) break;
                for (int32_t e = T.vptr[v]; e < T.vptr[v + 1]; ++e) {
                    const uint32_t* t = tri + 3 * (int64_t)T.vtri[e];
                    for (int k = 0; k < 3; ++k) {
                        const int32_t w = (int32_t)t[k];
                        if (pid[w] < 0) { pid[w] = np; queue.push_back(w); }
                    }
                }
            }
            // vertices queued but not placed go back to the pool
            for (size_t q = head; q < queue.size(); ++q)
                if (slot[queue[q]] < 0) pid[queue[q]] = -1;
            while (filled < PATCH) { T.patch_col.push_back(-1); ++filled; }
            ++np;
        }
        T.npatch = np;
        // patches never span loose parts (they grow through triangles); order
        // them by part so that every part owns a contiguous patch range
        {
            std::vector<int32_t> ppart(np), order(np), rank(np);
            for (int p = 0; p < np; ++p) ppart[p] = T.vpart[T.patch_col[(size_t)p * PATCH]];
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ppart[a] < ppart[b]; });
            for (int p = 0; p < np; ++p) rank[order[p]] = p;
            std::vector<int32_t> cols(T.patch_col.size());
            for (int p = 0; p < np; ++p)
                for (int s2 = 0; s2 < PATCH; ++s2) cols[(size_t)rank[p] * PATCH + s2] = T.patch_col[(size_t)p * PATCH + s2];
            T.patch_col.swap(cols);
            for (int64_t i = 0; i < N; ++i) pid[i] = rank[pid[i]];
            T.part_pptr.assign(T.nparts + 1, 0);
            for (int p = 0; p < np; ++p) T.part_pptr[ppart[order[p]] + 1]++;
            for (int c = 0; c < T.nparts; ++c) T.part_pptr[c + 1] += T.part_pptr[c];
        }
        T.patch_ptr.assign(np + 1, 0);
        for (int64_t j = 0; j < M; ++j) {
            int32_t ps[3] = {pid[tri[3 * j]], pid[tri[3 * j + 1]], pid[tri[3 * j + 2]]};
            for (int k = 0; k < 3; ++k) {
                bool first = true;
                for (int l = 0; l < k; ++l) if (ps[l] == ps[k]) first = false;
                if (first) T.patch_ptr[ps[k] + 1]++;
            }
        }
        for (int p = 0; p < np; ++p) T.patch_ptr[p + 1] += T.patch_ptr[p];
        const int64_t ne = T.patch_ptr[np];
        T.patch_tri.resize(ne);
        T.patch_slots.resize(ne);
        std::vector<int32_t> fill(T.patch_ptr.begin(), T.patch_ptr.end() - 1);
        for (int64_t j = 0; j < M; ++j) {
            const uint32_t* t = tri + 3 * j;
            int32_t ps[3] = {pid[t[0]], pid[t[1]], pid[t[2]]};
            for (int k = 0; k < 3; ++k) {
                bool first = true;
                for (int l = 0; l < k; ++l) if (ps[l] == ps[k]) first = false;
                if (!first) continue;
                uint32_t packed = 0;
                for (int l = 0; l < 3; ++l) {
                    const uint32_t sl = (ps[l] == ps[k]) ? (uint32_t)slot[t[l]] : 0xFFu;
                    packed |= sl << (8 * l);
                }
                // bit 24: the triangle's primary entry (the one of its first
                // vertex's patch; exactly one a triangle), which carries the
                // triangle's whole H phi row contribution
                if (k == 0) packed |= 1u << 24;
                const int32_t e = fill[ps[k]]++;
                T.patch_tri[e] = (uint32_t)j;
                T.patch_slots[e] = packed;
            }
        }
        T.redundancy = (double)ne / (double)M;
    }

    // dense problems and the assembly work table
    T.independent = independent;
    if (independent) {
        // part c must be the vertex range [v0_c, v0_c + n_c) with v0_c increasing
        std::vector<int64_t> first(T.nparts, -1), count(T.nparts, 0);
        for (int64_t i = 0; i < N; ++i) {
            const int c = T.vpart[i];
            if (first[c] < 0) first[c] = i;
            count[c]++;
        }
        int64_t next = 0;
        for (int c = 0; c < T.nparts; ++c) {
            BH_CHECK(first[c] == next, BEMHIP_ERR_BAD_MESH,
                     "independent parts must be contiguous vertex ranges in increasing order");
            next += count[c];
        }
        for (int64_t i = 1; i < N; ++i)
            BH_CHECK(T.vpart[i] >= T.vpart[i - 1], BEMHIP_ERR_BAD_MESH,
                     "independent parts must be contiguous vertex ranges in increasing order");
        for (int c = 0; c < T.nparts; ++c) {
            AsmProb pr{};
            pr.v0 = (int32_t)first[c];
            pr.n = (int32_t)count[c];
            pr.p0 = T.part_pptr[c];
            pr.p1 = T.part_pptr[c + 1];
            T.probs.push_back(pr);
        }
        // G blocks grouped by size (largest first, then part order), each
        // 256-B aligned in FP32 too: the problems of one size then sit at a
        // uniform stride, so that a batched factorisation addresses problem b
        // of a size class as first + b * stride
        std::vector<int> ord(T.nparts);
        for (int c = 0; c < T.nparts; ++c) ord[c] = c;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return T.probs[a].n > T.probs[b].n; });
        int64_t gofs = 0;
        for (int c : ord) {
            T.probs[c].gofs = gofs;
            gofs += gblock_stride(T.probs[c].n);
        }
    } else {
        AsmProb pr{};
        pr.v0 = 0;
        pr.n = (int32_t)N;
        pr.gofs = 0;
        pr.p0 = 0;
        pr.p1 = T.npatch;
        T.probs.push_back(pr);
    }
    // about ASM_WG_TARGET work-groups in total, shared in proportion to n^2
    // (many small items: the last round of work-groups leaves fewer compute
    // units idle)
    const double wg_target = ASM_WG_TARGET;
    double work = 0.0;
    for (const AsmProb& pr : T.probs) work += (double)pr.n * pr.n;
    for (size_t q = 0; q < T.probs.size(); ++q) {
        AsmProb& pr = T.probs[q];
        const int rb = (pr.n + ASM_ROWS - 1) / ASM_ROWS;
        const int npp = pr.p1 - pr.p0;
        const double want = std::max<double>(rb, wg_target * (double)pr.n * pr.n / work);
        int groups = (int)std::ceil(want / rb);
        groups = std::max(1, std::min(groups, npp));
        const int ppg = (npp + groups - 1) / groups;
        groups = (npp + ppg - 1) / ppg;
        pr.groups = groups;
        T.max_groups = std::max(T.max_groups, groups);
        for (int g = 0; g < groups; ++g)
            for (int b = 0; b < rb; ++b) {
                AsmItem it{};
                it.prob = (int32_t)q;
                it.row0 = b * ASM_ROWS;
                it.p0 = pr.p0 + g * ppg;
                it.p1 = std::min(pr.p1, it.p0 + ppg);
                it.g = g;
                T.items.push_back(it);
            }
    }
    return T;
}

template <class V> static void up(DevBuf& b, const std::vector<V>& v, hipStream_t s) {
    b.ensure(sizeof(V) * std::max<size_t>(1, v.size()));
    if (!v.empty()) BH_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice, s));
}

void DevTopology::upload(const Topology& t, hipStream_t s) {
    static std::atomic<uint64_t> uploads{0};
    gen = uploads.fetch_add(1) + 1;
    N = t.N;
    M = t.M;
    npatch = t.npatch;
    nparts = t.nparts;
    nentries = (int)t.patch_tri.size();
    nitems = (int)t.items.size();
    max_groups = t.max_groups;
    independent = t.independent;
    probs = t.probs;
    gsize = 0;
    for (const AsmProb& pr : probs) gsize = std::max<int64_t>(gsize, pr.gofs + (int64_t)pr.n * pr.n);
    up(d_probs, t.probs, s);
    up(d_items, t.items, s);
    up(tri, t.tri, s);
    up(patch_col, t.patch_col, s);
    up(patch_ptr, t.patch_ptr, s);
    up(patch_tri, t.patch_tri, s);
    up(patch_slots, t.patch_slots, s);
    up(vptr, t.vptr, s);
    up(vtri, t.vtri, s);
    up(vpart, t.vpart, s);
    up(part_tptr, t.part_tptr, s);
    up(part_tri, t.part_tri, s);
    up(part_vptr, t.part_vptr, s);
    up(part_vtx, t.part_vtx, s);
    BH_HIP(hipStreamSynchronize(s));
}

}  // namespace bemhip

// ColocSimPin::set_x_boundary (ColocSimPin.cpp:222-287): a vertex with
// |x| < 1e-5 is put on the wall (x = 0) and swapped with the last vertex not
// yet visited, which is looked at next; triangles are relabelled through the
// inverse permutation, and phi / psi (either may be null) move with their
// vertices.  The visiting order fixes the resulting vertex order, which the
// pinned columns and every per-vertex array depend on.
extern "C" int bemhip_pin_boundary(int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi, double* psi,
                                   int64_t* npin) {
    if (N <= 0 || M < 0 || !xyz || (M > 0 && !tri) || !npin) return BEMHIP_ERR_ARG;
    for (int64_t e = 0; e < 3 * M; ++e)
        if (tri[e] >= (uint64_t)N) return BEMHIP_ERR_BAD_MESH;
    std::vector<int64_t> perm(N), inv(N);
    std::iota(perm.begin(), perm.end(), 0);
    int64_t k = N, pinned = 0;
    for (int64_t i = 0; i < k; ++i) {
        if (std::fabs(xyz[3 * i]) < 1e-5) {
            xyz[3 * i] = 0.0;
            const int64_t l = k - 1;
            std::swap(perm[i], perm[l]);
            for (int d = 0; d < 3; ++d) std::swap(xyz[3 * i + d], xyz[3 * l + d]);
            if (phi) std::swap(phi[i], phi[l]);
            if (psi) std::swap(psi[i], psi[l]);
            --i;
            --k;
            ++pinned;
        }
    }
    for (int64_t i = 0; i < N; ++i) inv[perm[i]] = i;
    for (int64_t e = 0; e < 3 * M; ++e) tri[e] = (uint32_t)inv[tri[e]];
    *npin = pinned;
    return BEMHIP_OK;
}
