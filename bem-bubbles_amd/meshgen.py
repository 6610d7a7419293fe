"""Synthetic inputs for the per-stage BEM path: frequency-nu icospheres, real
spherical harmonics and the five benchmark configurations.

The reference builds its meshes with a frequency-subdivided geodesic
icosahedron (python_utils/icosphere/icosphere.py:34-68, 10*nu^2+2 vertices and
20*nu^2 outward-oriented faces) and reads them back as double precision
(Mesh/MeshIO.cpp).  This module generates the same family of meshes with its
own construction: the 20 faces of the icosahedron are found from its edge
graph, every face is split into a nu x nu barycentric lattice, lattice points
shared by neighbouring faces are merged through an exact integer key, and the
points are projected onto the sphere.  Vertex numbering therefore differs from
the reference's; every consumer takes the mesh as explicit arrays.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------

def _icosahedron():
    g = (1.0 + math.sqrt(5.0)) / 2.0
    pts = []
    for s1 in (-1.0, 1.0):
        for s2 in (-1.0, 1.0):
            pts.append((0.0, s1, s2 * g))
            pts.append((s1, s2 * g, 0.0))
            pts.append((s2 * g, 0.0, s1))
    v = np.array(pts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # edges join vertices at the minimal mutual distance
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
    emin = d[d > 1e-9].min()
    adj = np.abs(d - emin) < 1e-9
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            if not adj[i, j]:
                continue
            for k in range(j + 1, 12):
                if adj[i, k] and adj[j, k]:
                    a, b, c = v[i], v[j], v[k]
                    if np.dot(np.cross(b - a, c - a), a + b + c) > 0:
                        faces.append((i, j, k))
                    else:
                        faces.append((i, k, j))
    assert len(faces) == 20
    return v, np.array(faces, dtype=np.int64)


def icosphere(nu: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Geodesic icosphere with subdivision frequency nu.

    Returns (verts float64 (V,3), tris uint32 (F,3)) with V = 10 nu^2 + 2,
    F = 20 nu^2 and right-handed (outward) faces.
    """
    if nu < 1:
        raise ValueError("nu must be >= 1")
    base_v, base_f = _icosahedron()
    index = {}
    verts = []
    tris = []

    def vid(key, pos):
        j = index.get(key)
        if j is None:
            j = len(verts)
            index[key] = j
            verts.append(pos)
        return j

    for (A, B, C) in base_f:
        pa, pb, pc = base_v[A], base_v[B], base_v[C]
        lat = {}
        for i in range(nu + 1):
            for j in range(nu + 1 - i):
                wa, wb, wc = nu - i - j, i, j
                key = tuple(sorted((int(n), int(w)) for n, w in ((A, wa), (B, wb), (C, wc)) if w > 0))
                pos = (wa * pa + wb * pb + wc * pc) / nu
                lat[(i, j)] = vid(key, pos)
        for i in range(nu):
            for j in range(nu - i):
                tris.append((lat[(i, j)], lat[(i + 1, j)], lat[(i, j + 1)]))
                if i + j + 1 < nu:
                    tris.append((lat[(i + 1, j)], lat[(i + 1, j + 1)], lat[(i, j + 1)]))
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v * float(radius) + np.asarray(center, dtype=np.float64)[None, :]
    t = np.array(tris, dtype=np.uint32)
    assert v.shape[0] == 10 * nu * nu + 2 and t.shape[0] == 20 * nu * nu
    return np.ascontiguousarray(v), np.ascontiguousarray(t)


def hemisphere(nu: int, radius: float = 1.0):
    """Geodesic hemisphere x >= 0 on the wall x = 0: the four faces of the
    octahedron around +x, each split into nu^2 triangles and projected on the
    sphere.  Its open rim (4 nu vertices) lies exactly in the plane x = 0, so the
    wall image closes it into a sphere.  Returns (verts, tris), outward faces."""
    if nu < 1:
        raise ValueError("nu must be >= 1")
    apex = np.array([1.0, 0.0, 0.0])
    rim = [np.array(p, dtype=np.float64) for p in ((0, 1, 0), (0, 0, 1), (0, -1, 0), (0, 0, -1))]
    index, verts, tris = {}, [], []

    def vid(pos):
        key = tuple(np.round(pos * 2 * nu).astype(int))
        j = index.get(key)
        if j is None:
            j = len(verts)
            index[key] = j
            verts.append(pos)
        return j

    for f in range(4):
        A, B, Cc = apex, rim[f], rim[(f + 1) % 4]
        lat = {}
        for i in range(nu + 1):
            for j in range(nu + 1 - i):
                lat[(i, j)] = vid(((nu - i - j) * A + i * B + j * Cc) / nu)
        for i in range(nu):
            for j in range(nu - i):
                tris.append((lat[(i, j)], lat[(i + 1, j)], lat[(i, j + 1)]))
                if i + j + 1 < nu:
                    tris.append((lat[(i + 1, j)], lat[(i + 1, j + 1)], lat[(i, j + 1)]))
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v *= float(radius)
    t = np.array(tris, dtype=np.uint32)
    # outward orientation
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    t[flip] = t[flip][:, [0, 2, 1]]
    assert v.shape[0] == 2 * nu * nu + 2 * nu + 1 and t.shape[0] == 4 * nu * nu
    return np.ascontiguousarray(v), np.ascontiguousarray(t)


def join_meshes(parts):
    """Concatenate meshes (Mesh.cpp join_meshes semantics: offsets added)."""
    vs, ts, off = [], [], 0
    for v, t in parts:
        vs.append(v)
        ts.append(t.astype(np.int64) + off)
        off += v.shape[0]
    return np.ascontiguousarray(np.concatenate(vs)), np.ascontiguousarray(np.concatenate(ts).astype(np.uint32))


# --------------------------------------------------------------------------
# real spherical harmonics (orthonormal on the unit sphere)
# --------------------------------------------------------------------------

def real_sph_harm(lmax: int, xyz: np.ndarray, center=(0.0, 0.0, 0.0)):
    """All real orthonormal spherical harmonics up to degree lmax evaluated at
    the directions of xyz - center.  Returns dict {(l, m): values}."""
    p = np.asarray(xyz, dtype=np.float64) - np.asarray(center, dtype=np.float64)[None, :]
    r = np.linalg.norm(p, axis=1)
    ct = np.clip(p[:, 2] / r, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = np.arctan2(p[:, 1], p[:, 0])
    # fully normalised associated Legendre functions (4 pi normalisation / 4 pi)
    P = {}
    P[(0, 0)] = np.full_like(ct, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(1, lmax + 1):
        P[(m, m)] = -math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * st * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = math.sqrt(2.0 * m + 3.0) * ct * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            P[(l, m)] = a * (ct * P[(l - 1, m)] - b * P[(l - 2, m)])
    Y = {}
    for l in range(lmax + 1):
        Y[(l, 0)] = P[(l, 0)]
        for m in range(1, l + 1):
            Y[(l, m)] = math.sqrt(2.0) * P[(l, m)] * np.cos(m * ph)
            Y[(l, -m)] = math.sqrt(2.0) * P[(l, m)] * np.sin(m * ph)
    return Y


def random_sh_field(xyz, lmin, lmax, seed, center=(0.0, 0.0, 0.0)):
    """sum_{l=lmin..lmax, m} c_lm Y_lm with c_lm ~ N(0,1) drawn from `seed`."""
    rng = np.random.default_rng(seed)
    Y = real_sph_harm(lmax, xyz, center)
    out = np.zeros(np.asarray(xyz).shape[0])
    for l in range(lmin, lmax + 1):
        for m in range(-l, l + 1):
            out += rng.standard_normal() * Y[(l, m)]
    return out


# --------------------------------------------------------------------------
# benchmark configurations
# --------------------------------------------------------------------------

@dataclass
class Physics:
    p_inf: float = 1.0
    epsilon: float = 1.0
    sigma: float = 0.0
    gamma: float = 1.0
    field_amp: float = 0.0      # p_field(x, t) = amp * sin(2 pi freq t + phase)
    field_freq: float = 0.0
    field_phase: float = 0.0
    V_0: float | None = None    # None: total initial volume (Simulation.hpp:42)


@dataclass
class Case:
    name: str
    verts: np.ndarray
    tris: np.ndarray
    phi: np.ndarray
    physics: Physics = field(default_factory=Physics)
    dt: float = 1e-3
    steps: int = 1
    parts: list = field(default_factory=list)   # per-bubble (vertex offset, count)


def config_small(seed: int = 0) -> Case:
    """One sphere, nu=8 (642 vertices), phi = 1 + 0.01 * SH(l=1..4)."""
    v, t = icosphere(8)
    phi = 1.0 + 0.01 * random_sh_field(v, 1, 4, seed)
    return Case("sphere_nu8", v, t, phi, Physics(1.0, 1.0, 0.0, 1.0), dt=1e-3, steps=1,
                parts=[(0, v.shape[0])])


def config_two_bubbles(seed: int = 0, nu: int = 32) -> Case:
    """Two spheres (nu=32 each by default), R=1.0 at x=-1.5 and R=0.8 at x=+1.5,
    radial velocities +0.5 / -0.5 (phi = -R * Rdot), 1e-3 Gaussian jitter."""
    rng = np.random.default_rng(seed)
    v1, t1 = icosphere(nu, 1.0, (-1.5, 0.0, 0.0))
    v2, t2 = icosphere(nu, 0.8, (1.5, 0.0, 0.0))
    v, t = join_meshes([(v1, t1), (v2, t2)])
    v = v + 1e-3 * rng.standard_normal(v.shape)
    phi = np.concatenate([np.full(v1.shape[0], -1.0 * 0.5), np.full(v2.shape[0], -0.8 * -0.5)])
    return Case(f"two_spheres_nu{nu}", np.ascontiguousarray(v), t, phi, Physics(1.0, 1.0, 0.0, 1.0),
                dt=5e-4, steps=10, parts=[(0, v1.shape[0]), (v1.shape[0], v2.shape[0])])


def config_large(seed: int = 0, nu: int = 64) -> Case:
    """One sphere, nu=64 (N = 40962), phi = -1 + 0.02 * SH(l=1..8)."""
    v, t = icosphere(nu)
    phi = -1.0 + 0.02 * random_sh_field(v, 1, 8, seed)
    return Case(f"sphere_nu{nu}", v, t, phi, Physics(1.0, 1.0, 0.0, 1.0), dt=1e-3, steps=3,
                parts=[(0, v.shape[0])])


def config_adjoint(seed: int = 0, nu: int = 16) -> Case:
    """One sphere, nu=16 (2562 vertices) at rest (phi = 0) in
    p_inf(t) = 1 + 0.3 sin(pi t), 50 RK4 steps of 1e-2."""
    v, t = icosphere(nu)
    phi = np.zeros(v.shape[0])
    return Case(f"sphere_nu{nu}_forced", v, t, phi,
                Physics(1.0, 1.0, 0.0, 1.0, field_amp=0.3, field_freq=0.5),
                dt=1e-2, steps=50, parts=[(0, v.shape[0])])


def adjoint_target_phi(case: Case, seed: int = 1, rel: float = 0.01):
    """Initial potential of the target run: the case's phi plus a seeded
    perturbation of 1% (unit-variance real SH, degrees 0..4, times rel)."""
    return case.phi + rel * random_sh_field(case.verts, 0, 4, seed)


def config_batch(seed: int = 0, count: int = 256):
    """`count` independent spheres, nu in {4, 8, 16, 32}, R ~ U[0.5, 2],
    forcing amplitude ~ U[0.1, 0.5] at frequency 0.5."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        nu = int(rng.choice([4, 8, 16, 32]))
        R = float(rng.uniform(0.5, 2.0))
        A = float(rng.uniform(0.1, 0.5))
        v, t = icosphere(nu, R)
        out.append(Case(f"batch{b}_nu{nu}", v, t, np.zeros(v.shape[0]),
                        Physics(1.0, 1.0, 0.0, 1.0, field_amp=A, field_freq=0.5), dt=1e-2, steps=20,
                        parts=[(0, v.shape[0])]))
    return out


def join_batch(cases, spacing: float = 5.0):
    """Join independent single-bubble cases into one mesh whose loose parts are
    contiguous vertex ranges in case order, bubbles placed on a cubic grid
    `spacing` apart (their interaction is ignored by the batched path).
    Returns (verts, tris, phi, parts, field_amp, field_freq) with per-part
    forcing arrays."""
    side = max(1, int(np.ceil(len(cases) ** (1.0 / 3.0))))
    meshes, phis, parts, off = [], [], [], 0
    for b, c in enumerate(cases):
        g = np.array([b % side, (b // side) % side, b // (side * side)], dtype=np.float64)
        meshes.append((c.verts + spacing * g, c.tris))
        phis.append(c.phi)
        parts.append((off, c.verts.shape[0]))
        off += c.verts.shape[0]
    v, t = join_meshes(meshes)
    amp = np.array([c.physics.field_amp for c in cases])
    freq = np.array([c.physics.field_freq for c in cases])
    return v, t, np.concatenate(phis), parts, amp, freq
