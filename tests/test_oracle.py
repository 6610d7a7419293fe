"""Pins of the FP64 CPU oracle (CPU only).

The reference ships no tests or golden vectors for this path, so the oracle is
pinned by the reference's own quadrature tables (tests/golden), closed forms,
brute-force integration, invariants of the operators and the analytic sphere
and Rayleigh-Plesset solutions.  Each check is chosen so that a plausible
slip (wrong basis index, sign, missing Jacobian factor, wrong mirror image,
transposed storage) fails it.
"""

import json
import math
import os

import numpy as np
import pytest

import meshgen

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def fact(n):
    return math.factorial(n)


# ------------------------------------------------------------ quadrature --

def test_quadrature_matches_reference_tables(oracle_lib):
    ref = json.load(open(os.path.join(GOLD, "quadrature_reference.json")))
    q = oracle_lib.quadrature()
    gx = sorted(0.5 * (1.0 + np.array(ref["gauss_7_t"])))
    gw = np.array(ref["gauss_7_w"]) * 0.5
    gw = gw[np.argsort(0.5 * (1.0 + np.array(ref["gauss_7_t"])))]
    assert np.max(np.abs(q["gauss_x"] - gx)) < 2e-16
    assert np.max(np.abs(q["gauss_w"] - gw)) < 2e-16
    tx, ty = np.array(ref["quadrature_7_x"]), np.array(ref["quadrature_7_y"])
    tw = 0.5 * np.array(ref["quadrature_7_w_unhalved"])
    assert np.max(np.abs(q["tri_x"] - tx)) < 2e-16
    assert np.max(np.abs(q["tri_y"] - ty)) < 2e-16
    assert np.max(np.abs(q["tri_w"] - tw)) < 2e-16


def test_quadrature_exactness(oracle_lib):
    q = oracle_lib.quadrature()
    # degree-5 triangle rule: int_T x^a y^b = a! b! / (a+b+2)!
    for a in range(6):
        for b in range(6 - a):
            exact = fact(a) * fact(b) / fact(a + b + 2)
            got = np.sum(q["tri_w"] * q["tri_x"] ** a * q["tri_y"] ** b)
            assert abs(got - exact) < 1e-15, (a, b)
    # degree 6 is not integrated exactly (sanity that the check has teeth)
    assert abs(np.sum(q["tri_w"] * q["tri_x"] ** 6) - fact(6) / fact(8)) > 1e-8
    # 7-point Gauss-Legendre on (0,1): exact to degree 13
    for k in range(14):
        assert abs(np.sum(q["gauss_w"] * q["gauss_x"] ** k) - 1.0 / (k + 1)) < 1e-15, k


def test_quadrature19_matches_reference_table(oracle_lib):
    ref = json.load(open(os.path.join(GOLD, "quadrature_reference.json")))
    q = oracle_lib.quadrature19()
    assert np.max(np.abs(q["tri_x"] - np.array(ref["quadrature_19_x"]))) < 2e-16
    assert np.max(np.abs(q["tri_y"] - np.array(ref["quadrature_19_y"]))) < 2e-16
    assert np.max(np.abs(q["tri_w"] - 0.5 * np.array(ref["quadrature_19_w_unhalved"]))) < 2e-17
    # degree 9 (Dunavant): every monomial of degree <= 9 exactly, degree 10 not
    for a in range(10):
        for b in range(10 - a):
            exact = fact(a) * fact(b) / fact(a + b + 2)
            assert abs(np.sum(q["tri_w"] * q["tri_x"] ** a * q["tri_y"] ** b) - exact) < 1e-15, (a, b)
    assert abs(np.sum(q["tri_w"] * q["tri_x"] ** 10) - fact(10) / fact(12)) > 1e-9


# ---------------------------------------------------------------- panels --

def _sub_rule(a, b, c, level):
    """Points/weights of the 7-point rule on 4^level sub-triangles (numpy, independent)."""
    s = math.sqrt(15.0)
    A1, B1, A2, B2 = (9 + 2 * s) / 21, (6 - s) / 21, (9 - 2 * s) / 21, (6 + s) / 21
    bary = np.array([[1 / 3, 1 / 3, 1 / 3], [A1, B1, B1], [B1, A1, B1], [B1, B1, A1],
                     [A2, B2, B2], [B2, A2, B2], [B2, B2, A2]])
    wts = np.array([9 / 40] + [(155 - s) / 1200] * 3 + [(155 + s) / 1200] * 3)
    tris = [(np.array(a, float), np.array(b, float), np.array(c, float))]
    for _ in range(level):
        nt = []
        for (p, q_, r) in tris:
            pq, qr, rp = (p + q_) / 2, (q_ + r) / 2, (r + p) / 2
            nt += [(p, pq, rp), (pq, q_, qr), (rp, qr, r), (pq, qr, rp)]
        tris = nt
    P, W = [], []
    for (p, q_, r) in tris:
        area = 0.5 * np.linalg.norm(np.cross(q_ - p, r - p))
        P.append(bary @ np.stack([p, q_, r]))
        W.append(wts * area)
    return np.concatenate(P), np.concatenate(W)


def _barycentric(P, a, b, c):
    T = np.stack([np.array(b) - np.array(a), np.array(c) - np.array(a)], axis=1)
    lam = np.linalg.lstsq(T, (P - np.array(a)).T, rcond=None)[0].T
    return np.stack([1 - lam[:, 0] - lam[:, 1], lam[:, 0], lam[:, 1]], axis=1)


def test_far_panel_against_brute_force(oracle_lib):
    a, b, c = [0.1, 0.0, 0.05], [0.9, 0.2, -0.1], [0.3, 0.7, 0.2]
    x = [0.5, -1.2, 1.6]
    g, h = oracle_lib.panel(x, a, b, c, -1, False)
    P, W = _sub_rule(a, b, c, 5)
    phi = _barycentric(P, a, b, c)
    z = P - np.array(x)
    r = np.linalg.norm(z, axis=1)
    n = np.cross(np.array(b) - a, np.array(c) - np.array(b))
    n /= np.linalg.norm(n)
    gb = (phi * (W / r)[:, None]).sum(0)
    hb = (phi * (W * (-(z @ n)) / r ** 3)[:, None]).sum(0)
    assert np.allclose(g, gb, rtol=2e-5, atol=0)
    assert np.allclose(h, hb, rtol=2e-5, atol=0)
    # sum_k H_k = -(solid angle of T seen from x), Van Oosterom-Strackee
    R1, R2, R3 = (np.array(v) - np.array(x) for v in (a, b, c))
    r1, r2, r3 = map(np.linalg.norm, (R1, R2, R3))
    num = R1 @ np.cross(R2, R3)
    den = r1 * r2 * r3 + (R1 @ R2) * r3 + (R1 @ R3) * r2 + (R2 @ R3) * r1
    omega = 2 * math.atan2(num, den)
    assert abs(h.sum() + omega) < 1e-5 * abs(omega)


def test_self_panel_against_duffy(oracle_lib):
    a, b, c = np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.1, 0.0]), np.array([0.3, 0.8, 0.1])
    for sv, (A, B, Cc) in enumerate([(a, b, c), (b, c, a), (c, a, b)]):
        # collocation point is vertex `sv` of the ordered triangle (a, b, c)
        g, h = oracle_lib.panel([a, b, c][sv], a, b, c, sv, False)
        # Duffy: y = A + s (B - A) + s t (C - B), dA = J s ds dt
        ab, bc = B - A, Cc - B
        J = np.linalg.norm(np.cross(ab, bc))
        xs, ws = np.polynomial.legendre.leggauss(60)
        s, ws_ = 0.5 * (xs + 1), 0.5 * ws
        S, T = np.meshgrid(s, s, indexing="ij")
        Wt = np.outer(ws_, ws_)
        d = np.linalg.norm(ab[None, None, :] + T[..., None] * bc[None, None, :], axis=2)
        basis = [1 - S, S - S * T, S * T]   # vertices A, B, C of the reordered triangle
        vals = np.array([(Wt * J * f / d).sum() for f in basis])
        # map reordered vertex k back to the triangle's vertex (k + sv) % 3
        want = np.zeros(3)
        for k in range(3):
            want[(k + sv) % 3] = vals[k]
        assert np.allclose(g, want, rtol=1e-7, atol=0), (sv, g, want)   # 7-point polar rule error ~1e-8
        assert np.all(h == 0.0)


# -------------------------------------------------------------- assembly --

def test_h_rows_sum_to_minus_four_pi(oracle_lib):
    v, t = meshgen.icosphere(4, 1.3, (0.2, -0.1, 0.3))
    G, H = oracle_lib.assemble(v, t)
    assert np.max(np.abs(H.sum(axis=1) + 4 * math.pi)) < 1e-12
    # G is stored column-major: column j collects the basis of vertex j
    G2, _ = oracle_lib.assemble(v[::-1].copy(), (len(v) - 1 - t.astype(np.int64)).astype(np.uint32))
    assert np.allclose(G2[::-1, ::-1], G, rtol=1e-13, atol=1e-15)


def _psi_err(oracle_lib, nu, field_fn, solver="lu"):
    v, t = meshgen.icosphere(nu, 1.0)
    p = oracle_lib.make_params(verts=v, tris=t, solver=solver)
    phi, psi_exact = field_fn(v)
    psi = oracle_lib.compute_psi(p, v, t, phi)
    return np.max(np.abs(psi - psi_exact)) / np.max(np.abs(psi_exact))


def test_sphere_analytic_solutions_converge(oracle_lib):
    uniform = lambda v: (np.full(len(v), 2.0), np.full(len(v), -2.0))   # psi = -phi0 / R
    e4, e8 = _psi_err(oracle_lib, 4, uniform), _psi_err(oracle_lib, 8, uniform)
    assert e8 < 0.01 and 3.0 < e4 / e8 < 5.5          # O(h^2)
    for (l, m) in [(1, 0), (2, -1), (3, 2)]:
        ylm = lambda v, l=l, m=m: (meshgen.real_sph_harm(l, v)[(l, m)], -(l + 1) * meshgen.real_sph_harm(l, v)[(l, m)])
        e4, e8 = _psi_err(oracle_lib, 4, ylm), _psi_err(oracle_lib, 8, ylm)
        assert e8 < e4 / 2.5, (l, m, e4, e8)


def test_mirror_equals_doubled_mesh(oracle_lib):
    v, t = meshgen.icosphere(3, 0.7, (1.1, 0.2, -0.1))
    N = len(v)
    vm = v.copy()
    vm[:, 0] *= -1.0
    tm = t[:, [0, 2, 1]].astype(np.int64) + N   # reversed orientation keeps normals outward
    vf = np.concatenate([v, vm])
    tf = np.concatenate([t.astype(np.int64), tm]).astype(np.uint32)
    Gm, Hm = oracle_lib.assemble(v, t, mirror=True)
    Gf, Hf = oracle_lib.assemble(vf, tf, mirror=False)
    assert np.allclose(Gm, Gf[:N, :N] + Gf[:N, N:], rtol=1e-12, atol=1e-14)
    assert np.allclose(Hm, Hf[:N, :N] + Hf[:N, N:], rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------- dense solves --

def test_lu_against_lapack(oracle_lib):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((300, 300))          # needs real row interchanges
    b = rng.standard_normal(300)
    LU, piv, info = oracle_lib.lu_factor(A)
    assert info == 0
    assert np.any(piv != np.arange(300))
    x = oracle_lib.lu_solve(LU, piv, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # P A = L U
    L = np.tril(LU, -1) + np.eye(300)
    U = np.triu(LU)
    PA = A.copy()
    for k, p in enumerate(piv):
        PA[[k, p]] = PA[[p, k]]
    assert np.max(np.abs(PA - L @ U)) < 1e-12 * np.max(np.abs(A)) * 300
    v, t = meshgen.icosphere(6)
    G, _ = oracle_lib.assemble(v, t)
    bb = rng.standard_normal(len(v))
    assert np.allclose(oracle_lib.solve(G, bb), np.linalg.solve(G, bb), rtol=1e-10, atol=1e-13)


def test_bicgstab_against_lapack(oracle_lib):
    v, t = meshgen.icosphere(6)
    G, H = oracle_lib.assemble(v, t)
    b = H @ (1.0 + 0.1 * meshgen.random_sh_field(v, 1, 3, 1))
    x, it, err = oracle_lib.bicgstab(G, b)
    assert 0 < it < 2 * len(v) and err < 1e-15
    ref = np.linalg.solve(G, b)
    assert np.max(np.abs(x - ref)) < 1e-11 * np.max(np.abs(ref))


# -------------------------------------------------------------- geometry --

def test_geometry_on_sphere(oracle_lib):
    from scipy.spatial import ConvexHull
    R, ctr = 1.7, np.array([0.3, -0.2, 0.5])
    v, t = meshgen.icosphere(6, R, ctr)
    vn = oracle_lib.vertex_normals(v, t)
    radial = (v - ctr) / np.linalg.norm(v - ctr, axis=1, keepdims=True)
    assert np.max(np.abs(vn - radial)) < 1e-13     # Max 1999 weights are exact on spheres
    k, gk = oracle_lib.curvatures(v, t)
    assert np.max(np.abs(k - 1.0 / R)) < 1e-9
    assert np.max(np.abs(gk - 1.0 / R ** 2)) < 1e-8
    vol = oracle_lib.part_volumes(v, t)
    assert abs(vol[0] - ConvexHull(v).volume) < 1e-12 * vol[0]
    tn = oracle_lib.triangle_normals(v, t)
    assert np.all(np.einsum("ij,ij->i", tn, v[t[:, 0]] - ctr) > 0)
    # tangential gradient of a linear field is its projection on each triangle
    gvec = np.array([0.3, -1.1, 0.7])
    tg = oracle_lib.tangent_gradients(v, t, v @ gvec)
    want = gvec[None, :] - (tn @ gvec)[:, None] * tn
    assert np.max(np.abs(tg - want)) < 1e-11


def test_components_and_volumes(oracle_lib):
    c = meshgen.config_two_bubbles(nu=4)
    comp, n = oracle_lib.components(len(c.verts), c.tris)
    assert n == 2 and comp[0] == 0 and comp[-1] == 1
    vol = oracle_lib.part_volumes(c.verts, c.tris)
    assert abs(vol[0] / vol[1] - (1.0 / 0.8) ** 3) < 0.02


# ---------------------------------------------------------------- RK4 ----

def test_rk4_follows_rayleigh_plesset(oracle_lib):
    from scipy.integrate import solve_ivp
    nu, Rdot0, dt, steps = 8, 0.1, 1e-2, 10
    v, t = meshgen.icosphere(nu)
    ph = meshgen.Physics(p_inf=1.0, epsilon=1.0, sigma=0.0, gamma=1.0)
    p = oracle_lib.make_params(ph, verts=v, tris=t)
    sim = oracle_lib.Sim(p, v, t, np.full(len(v), -1.0 * Rdot0))
    V0 = oracle_lib.part_volumes(v, t)[0]
    radii = []
    for _ in range(steps):
        sim.rk4(dt, True)
        radii.append((oracle_lib.part_volumes(sim.x, t)[0] / V0) ** (1.0 / 3.0))
    # R R'' + 3/2 R'^2 = -p_inf + eps (R0/R)^(3 gamma), R(0) = 1, R'(0) = Rdot0
    f = lambda _t, y: [y[1], (-1.5 * y[1] ** 2 - 1.0 + (1.0 / y[0]) ** 3) / y[0]]
    sol = solve_ivp(f, (0, dt * steps), [1.0, Rdot0], rtol=1e-12, atol=1e-14,
                    t_eval=dt * np.arange(1, steps + 1))
    assert np.max(np.abs(np.array(radii) - sol.y[0])) < 2e-4
    assert radii[-1] > 1.005          # the bubble grows: sign of the coupling
    # psi after a step is the averaged velocity projected on the new normals
    assert np.allclose(sim.psi, sim.psi.mean(), rtol=1e-2)


# ------------------------------------------------- potential off the surface --

def test_exterior_panel_against_brute_force(oracle_lib):
    a, b, c = [0.1, 0.0, 0.05], [0.9, 0.2, -0.1], [0.3, 0.7, 0.2]
    # (point, tolerance on G, on H): the rule's own error grows as y nears the panel
    for y, tg, th in (([0.5, -1.2, 1.6], 1e-10, 1e-9), ([0.4, 0.3, 0.6], 1e-5, 1e-4)):
        g, h = oracle_lib.exterior_panel(y, a, b, c)
        P, W = _sub_rule(a, b, c, 5)
        phi = _barycentric(P, a, b, c)
        z = P - np.array(y)
        r = np.linalg.norm(z, axis=1)
        n = np.cross(np.array(b) - a, np.array(c) - np.array(b))
        n /= np.linalg.norm(n)
        gb = (phi * (W / r)[:, None]).sum(0)
        hb = (phi * (W * (-(z @ n)) / r ** 3)[:, None]).sum(0)
        assert np.allclose(g, gb, rtol=tg, atol=0), (y, g, gb)
        assert np.allclose(h, hb, rtol=th, atol=0), (y, h, hb)
        # the degree-9 rule beats the degree-5 rule of the matrices by far
        g7, h7 = oracle_lib.panel(y, a, b, c, -1, False)
        assert np.max(np.abs(g - gb)) * 50 < np.max(np.abs(g7 - gb))
        assert np.max(np.abs(h - hb)) * 50 < np.max(np.abs(h7 - hb))


def _sphere_points(rng, count, rmin, rmax):
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * rng.uniform(rmin, rmax, size=(count, 1))


def test_exterior_potential_green_identities_on_sphere(oracle_lib):
    """Green's representation on the unit sphere: with n pointing out of the
    sphere, 1/(4 pi) int [phi d_n(1/r) - psi / r] dS equals u(y) outside for a
    decaying exterior field u and -u(y) inside for a field harmonic inside,
    and vanishes on the other side; the discrete values converge as O(h^2)."""
    rng = np.random.default_rng(5)
    out_pts = _sphere_points(rng, 12, 1.5, 3.0)
    in_pts = _sphere_points(rng, 12, 0.0, 0.5)
    errs = []
    for nu in (4, 8):
        v, t = meshgen.icosphere(nu, 1.0)
        one = np.ones(len(v))
        # u = 1/|y|: phi = 1, psi = d_r u = -1
        e_out = oracle_lib.exterior_pot(v, t, one, -one, out_pts) - 1.0 / np.linalg.norm(out_pts, axis=1)
        e_in0 = oracle_lib.exterior_pot(v, t, one, -one, in_pts)
        # constant phi, psi = 0: 0 outside, -1 inside (solid angle)
        c_out = oracle_lib.exterior_pot(v, t, one, 0 * one, out_pts)
        c_in = oracle_lib.exterior_pot(v, t, one, 0 * one, in_pts) + 1.0
        # u = z + 2 x y (harmonic inside): phi = u, psi = grad u . n; -u inside, 0 outside
        u = v[:, 2] + 2 * v[:, 0] * v[:, 1]
        du = v[:, 2] + 4 * v[:, 0] * v[:, 1]                # r d_r of a degree-1 plus a degree-2 term
        l_in = oracle_lib.exterior_pot(v, t, u, du, in_pts) + (in_pts[:, 2] + 2 * in_pts[:, 0] * in_pts[:, 1])
        l_out = oracle_lib.exterior_pot(v, t, u, du, out_pts)
        errs.append([np.max(np.abs(e)) for e in (e_out, e_in0, c_out, c_in, l_in, l_out)])
    e4, e8 = np.array(errs[0]), np.array(errs[1])
    assert np.all(e8 < 5e-3), e8
    assert np.all(e4 / e8 > 3.0), (e4, e8)                   # O(h^2)
    # the flat panels' solid angles close exactly: only the rule's error is left
    assert np.all(e4[2:4] < 1e-9) and np.all(e8[2:4] < 1e-12), (e4, e8)


def test_exterior_potential_is_sum_of_panels(oracle_lib):
    v, t = meshgen.icosphere(2, 1.1, (0.1, 0.0, -0.2))
    rng = np.random.default_rng(2)
    phi, psi = rng.normal(size=len(v)), rng.normal(size=len(v))
    y = np.array([[1.7, -0.4, 0.9]])
    want = 0.0
    for (ia, ib, ic) in t:
        g, h = oracle_lib.exterior_panel(y[0], v[ia], v[ib], v[ic])
        want += (-g @ psi[[ia, ib, ic]] + h @ phi[[ia, ib, ic]]) / (4 * math.pi)
    got = oracle_lib.exterior_pot(v, t, phi, psi, y, nthreads=1)
    assert abs(got[0] - want) < 1e-14 * max(1.0, abs(want))


def test_blocked_lu_matches_gepp(oracle_lib):
    """The CPU baseline's cache-blocked GEPP chooses the same pivots as the
    oracle's GEPP (PartialPivLU's first maximum) and reproduces LAPACK; a
    sample of the first panels leaves the rest of the matrix updated by
    exactly those panels."""
    rng = np.random.default_rng(4)
    n = 900
    A = rng.standard_normal((n, n))
    L1, p1, _ = oracle_lib.lu_factor(A)
    L2, p2, info = oracle_lib.lu_factor_blocked(A, nb=128)
    assert info == 0 and np.array_equal(p1, p2)
    assert np.max(np.abs(L1 - L2)) < 1e-10 * np.max(np.abs(L1))
    b = rng.standard_normal(n)
    x = oracle_lib.lu_solve(L2, p2, b)
    assert np.max(np.abs(x - np.linalg.solve(A, b))) < 1e-9 * np.max(np.abs(x))
    L3, p3, _ = oracle_lib.lu_factor_blocked(A, nb=128, max_panels=2)
    assert np.array_equal(p3[:256], p2[:256])
    # rows below 256 are interchanged by the later panels: the leading block only
    assert np.max(np.abs(L3[:256, :256] - L2[:256, :256])) < 1e-10 * np.max(np.abs(L2))
