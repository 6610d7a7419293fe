"""State files (binary PLY, MeshIO.cpp) and the pinned contact line
(ColocSimPin.cpp) on the CPU: the library's host-side reader / writer and
vertex reordering, and the oracle's pinned formulation.

The PLY checks hold the library to the byte layout the reference writes
(header text, 40-byte double records, 13-byte faces) and read the reference's
own input file tests/golden/init-pinned.ply (a data file of the reference,
init_conditions/init-pinned.ply: 2121 float vertices, 4160 faces), parsed
independently with numpy.  The pinned formulation is pinned by the original
BEM equations G psi = H phi holding for the mixed unknowns, by the analytic
uniform-potential solution on a hemisphere on the wall, and by the reference's
vertex reordering applied by two independent implementations."""

import math
import os

import numpy as np
import pytest

import bemhip
import meshgen

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PINNED_PLY = os.path.join(GOLD, "init-pinned.ply")


def _numpy_ply(path):
    """independent parse of a binary PLY with float or double x y z [values] and uchar/uint faces"""
    b = open(path, "rb").read()
    end = b.index(b"end_header\n") + len(b"end_header\n")
    head = b[:end].decode().splitlines()
    nv = nf = 0
    vt, names = [], []
    cur = None
    for ln in head:
        w = ln.split()
        if w[:2] == ["element", "vertex"]:
            nv, cur = int(w[2]), "v"
        elif w[:2] == ["element", "face"]:
            nf, cur = int(w[2]), "f"
        elif w[0] == "property" and cur == "v":
            vt.append({"float": "<f4", "double": "<f8"}[w[1]])
            names.append(w[2])
    rec = np.dtype([(n, t) for n, t in zip(names, vt)])
    verts = np.frombuffer(b, dtype=rec, count=nv, offset=end)
    faces = np.frombuffer(b, dtype=np.dtype([("n", "u1"), ("i", "<u4", (3,))]), count=nf,
                          offset=end + rec.itemsize * nv)
    return verts, faces, b[:end]


# ------------------------------------------------------------------ PLY --

def test_read_reference_input_file():
    v, t, phi, psi = bemhip.read_ply(PINNED_PLY)
    rv, rf, _ = _numpy_ply(PINNED_PLY)
    assert v.shape == (2121, 3) and t.shape == (4160, 3) and phi is None and psi is None
    assert np.all(rf["n"] == 3)
    want = np.stack([rv["x"], rv["y"], rv["z"]], axis=1).astype(np.float64)
    assert np.array_equal(v, want)          # float -> double is exact
    assert np.array_equal(t, rf["i"])


def test_double_state_file_round_trip_and_layout(tmp_path):
    v, t = meshgen.icosphere(3, 1.3, (0.2, 0.0, -0.1))
    rng = np.random.default_rng(0)
    phi, psi = rng.normal(size=len(v)), rng.normal(size=len(v))
    f = str(tmp_path / "state.ply")
    bemhip.write_ply(f, v, t, phi, psi, precision=8)
    # export_ply_double's header (MeshIO.cpp:306-317) and record sizes
    head = ("ply\nformat binary_little_endian 1.0\ncomment Output from BEM simulation\n"
            f"element vertex {len(v)}\nproperty double x\nproperty double y\nproperty double z\n"
            "property double phi\nproperty double psi\n"
            f"element face {len(t)}\nproperty list uchar uint vertex_indices\nend_header\n").encode()
    raw = open(f, "rb").read()
    assert raw[:len(head)] == head
    assert len(raw) == len(head) + 40 * len(v) + 13 * len(t)
    rv, rf, _ = _numpy_ply(f)
    assert np.array_equal(np.stack([rv["x"], rv["y"], rv["z"]], 1), v)
    assert np.array_equal(rv["phi"], phi) and np.array_equal(rv["psi"], psi)
    assert np.all(rf["n"] == 3) and np.array_equal(rf["i"], t)
    v2, t2, phi2, psi2 = bemhip.read_ply(f)
    assert np.array_equal(v2, v) and np.array_equal(t2, t)
    assert np.array_equal(phi2, phi) and np.array_equal(psi2, psi)


def test_float_and_one_value_files(tmp_path):
    v, t = meshgen.icosphere(2)
    phi = np.linspace(-1.0, 1.0, len(v)) * math.pi
    f4 = str(tmp_path / "f.ply")
    bemhip.write_ply(f4, v, t, phi, 2 * phi, precision=4)          # Simulation::export_mesh's layout
    assert open(f4, "rb").read().count(b"property float") == 5
    v2, t2, phi2, psi2 = bemhip.read_ply(f4)
    assert np.array_equal(v2, v.astype(np.float32).astype(np.float64))
    assert np.array_equal(phi2, phi.astype(np.float32).astype(np.float64))
    assert np.array_equal(psi2, (2 * phi).astype(np.float32).astype(np.float64))
    f1 = str(tmp_path / "w.ply")
    bemhip.write_ply(f1, v, t, phi, None, precision=8)              # one value named w (MeshIO.cpp:227-261)
    assert b"property double w\n" in open(f1, "rb").read()
    _, _, p1, s1 = bemhip.read_ply(f1)
    assert np.array_equal(p1, phi) and s1 is None
    f0 = str(tmp_path / "m.ply")
    bemhip.write_ply(f0, v, t, precision=4)                         # export_ply (MeshIO.cpp:67-100)
    v0, t0, p0, s0 = bemhip.read_ply(f0)
    assert p0 is None and s0 is None and np.array_equal(t0, t)


def test_extra_vertex_and_face_properties_are_skipped(tmp_path):
    """Blender-style x y z nx ny nz s t + a face flag: as in import_ply, the 4th
    and 5th float are taken as phi and psi and the rest of each record is
    skipped (MeshIO.cpp:768-843, 850-852)."""
    v, t = meshgen.icosphere(1)
    n = len(v)
    rec = np.zeros(n, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("nx", "<f4"), ("ny", "<f4"),
                             ("nz", "<f4"), ("s", "<f4"), ("t", "<f4")])
    for k, c in enumerate("xyz"):
        rec[c] = v[:, k]
    rec["nx"], rec["ny"] = 0.5, -0.25
    fr = np.zeros(len(t), dtype=[("n", "u1"), ("i", "<i4", (3,)), ("flag", "u1")])
    fr["n"], fr["i"], fr["flag"] = 3, t.astype(np.int32), 7
    head = ("ply\nformat binary_little_endian 1.0\ncomment x\n"
            f"element vertex {n}\n" + "".join(f"property float {c}\n" for c in ("x", "y", "z", "nx", "ny", "nz", "s", "t"))
            + f"element face {len(t)}\nproperty list uchar int vertex_indices\nproperty uchar flag\nend_header\n")
    f = str(tmp_path / "blender.ply")
    open(f, "wb").write(head.encode() + rec.tobytes() + fr.tobytes())
    v2, t2, phi, psi = bemhip.read_ply(f)
    assert np.array_equal(v2, v.astype(np.float32).astype(np.float64)) and np.array_equal(t2, t)
    assert np.all(phi == 0.5) and np.all(psi == -0.25)


def test_bad_files_are_reported(tmp_path):
    v, t = meshgen.icosphere(1)
    good = str(tmp_path / "g.ply")
    bemhip.write_ply(good, v, t, precision=8)
    raw = open(good, "rb").read()
    cases = {
        "missing": None,
        "ascii": raw.replace(b"binary_little_endian", b"ascii"),
        "truncated": raw[:-5],
        "quad": raw[:raw.index(b"end_header\n") + 11 + 24 * len(v)] + b"\x04" + raw[raw.index(b"end_header\n") + 12 + 24 * len(v):],
        "index": raw[:-4] + np.uint32(len(v)).tobytes(),
    }
    for name, data in cases.items():
        f = str(tmp_path / f"{name}.ply")
        if data is not None:
            open(f, "wb").write(data)
        with pytest.raises(bemhip.BemHipError) as e:
            bemhip.read_ply(f)
        assert e.value.code == bemhip.ERR_IO, name
    with pytest.raises(bemhip.BemHipError):
        bemhip.write_ply(str(tmp_path / "x.ply"), v, t, precision=2)


# --------------------------------------------------------------- pinning --

def test_pin_boundary_library_equals_oracle_on_reference_input(oracle_lib):
    v, t, _, _ = bemhip.read_ply(PINNED_PLY)
    phi = np.arange(len(v), dtype=np.float64)
    psi = -phi
    a = bemhip.pin_boundary(v, t, phi, psi)
    b = oracle_lib.pin_boundary(v, t, phi, psi)
    assert a[4] == b[4] == 80
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    v2, t2, phi2, psi2, npin = a
    # the pinned set is {|x| < 1e-5}, at the end, on the wall
    assert np.array_equal(np.sort(phi2[-npin:]).astype(int), np.flatnonzero(np.abs(v[:, 0]) < 1e-5))
    assert np.all(v2[-npin:, 0] == 0.0) and np.all(np.abs(v2[:-npin, 0]) >= 1e-5)
    # a relabelling: every vertex carries its own values, triangles keep their corners
    old = phi2.astype(int)
    snapped = v.copy()
    snapped[np.abs(snapped[:, 0]) < 1e-5, 0] = 0.0
    assert np.array_equal(v2, snapped[old]) and np.array_equal(psi2, -phi2)
    assert np.array_equal(old[t2], t)


def test_pinned_solve_satisfies_the_bem_equations(oracle_lib):
    """The swapped system's unknowns are psi (free) and phi (pinned): with them
    the original G psi = H phi must hold; a sign slip in the column swap fails
    this by O(1)."""
    v, t = meshgen.hemisphere(6)
    v, t, _, _, npin = oracle_lib.pin_boundary(v, t)
    N = len(v)
    rng = np.random.default_rng(1)
    pot = rng.normal(size=N)          # phi (free), prescribed psi (pinned)
    p = oracle_lib.make_params(verts=v, tris=t, mirror=True, n_pin=npin)
    x = oracle_lib.compute_psi(p, v, t, pot)
    G, H = oracle_lib.assemble(v, t, mirror=True)
    phi, psi = pot.copy(), x.copy()
    phi[N - npin:], psi[N - npin:] = x[N - npin:], pot[N - npin:]
    assert np.max(np.abs(G @ psi - H @ phi)) < 1e-11 * np.max(np.abs(H @ phi))


def test_pinned_hemisphere_uniform_potential(oracle_lib):
    """A hemisphere of radius 1 on the wall (closed by its image) at uniform
    phi = 1: psi = -1 is exact; prescribing it on the rim, the solve returns
    psi = -1 on the free vertices (O(h^2)) and phi = 1 on the rim."""
    errs = []
    for nu in (4, 8, 16):
        v, t = meshgen.hemisphere(nu)
        v, t, _, _, npin = oracle_lib.pin_boundary(v, t)
        assert npin == 4 * nu
        N = len(v)
        pot = np.ones(N)
        pot[N - npin:] = -1.0
        p = oracle_lib.make_params(verts=v, tris=t, mirror=True, n_pin=npin)
        x = oracle_lib.compute_psi(p, v, t, pot)
        errs.append((np.max(np.abs(x[:N - npin] + 1.0)), np.max(np.abs(x[N - npin:] - 1.0))))
    e = np.array(errs)
    assert e[-1, 0] < 5e-3 and e[-1, 1] < 5e-3, e
    assert np.all(e[:-1, 0] / e[1:, 0] > 3.5), e          # O(h^2) for psi
    assert np.all(e[:-1, 1] / e[1:, 1] > 1.8), e          # the rim converges too


def test_drift_law(oracle_lib):
    """ColocSimPin.cpp:126-136: beta = 0.2, theta_eq = 130 deg, c_nl = 0.01; the
    line advances (positive drift) for contact angles above equilibrium and
    recedes below it, and the nonlinear term blows up as cos(theta) -> +-1."""
    c_eq = math.cos(130.0 / 180.0 * math.pi)
    d = oracle_lib.drift_velocity
    assert abs(d(c_eq) - 0.01 * c_eq ** 3 / (c_eq ** 2 - 1.0)) < 1e-15
    assert abs(d(0.0) - 0.2 * c_eq) < 1e-15
    # the root sits between theta_eq and 90 deg, shifted by the nonlinear term
    lo, hi = c_eq, 0.0
    assert d(lo) > 0.0 > d(hi)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if d(mid) > 0.0 else (lo, mid)
    assert c_eq < lo < c_eq + 0.05
    assert d(0.999) < -1.0 and d(-0.999) > 1.0


def test_pinned_euler_keeps_the_solved_phi(oracle_lib):
    """position_t rewrites the pinned entries of its potential (ColocSimPin.cpp:
    146-148, 186): after an Euler step phi_pinned = (solved phi) + dt dphi/dt,
    not the old phi; pinned vertices stay on the wall."""
    v, t = meshgen.hemisphere(5)
    v, t, _, _, npin = oracle_lib.pin_boundary(v, t)
    N = len(v)
    rng = np.random.default_rng(3)
    phi0 = 0.1 * rng.normal(size=N)
    psi0 = 0.1 * rng.normal(size=N)
    p = oracle_lib.make_params(verts=v, tris=t, mirror=True, n_pin=npin)
    sim = oracle_lib.Sim(p, v, t, phi0)
    sim.psi = psi0.copy()
    dt = 1e-3
    sim.euler(dt, True)
    pot = phi0.copy()
    pot[N - npin:] = psi0[N - npin:]
    x = oracle_lib.compute_psi(p, v, t, pot)              # the pinned solve at the step start
    assert np.all(sim.x[N - npin:, 0] == 0.0)             # the drift is along the wall
    vel = (sim.x - v) / dt                                # x_new = x + dt vel
    assert np.max(np.abs(vel[N - npin:])) > 0.0
    dpot = oracle_lib.pot_t_multi(p, v, t, vel, 0.0)
    want = phi0 + dt * dpot
    want[N - npin:] = x[N - npin:] + dt * dpot[N - npin:]
    assert np.max(np.abs(sim.phi - want)) < 1e-12
    assert np.max(np.abs(x[N - npin:] - phi0[N - npin:])) > 1e-3    # the rewrite matters here
