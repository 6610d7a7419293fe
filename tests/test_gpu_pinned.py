"""Pinned contact line (ColocSimPin) and PLY state files on the GPU: the
device path through the C-ABI against the FP64 oracle on the reference's own
input mesh (init_conditions/init-pinned.ply, 2121 vertices, 80 pinned) and on
geodesic hemispheres, plus export / restart through binary PLY."""

import os
import sys

import numpy as np
import pytest

import bemhip
import meshgen

pytestmark = pytest.mark.gpu

PINNED_PLY = os.path.join(os.path.dirname(__file__), "golden", "init-pinned.ply")


class Phys:
    p_inf, epsilon, sigma, gamma = 1.0, 1.2, 0.0, 1.4


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu(ctx, v, t, solver, phi=None, psi=None, **kw):
    ph = Phys
    return bemhip.ColocSimPin(ctx, v, t, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, solver=solver, phi=phi, psi=psi,
                              **kw)


def _oracle(oracle_lib, v, t, phi=None, psi=None):
    v2, t2, ph2, ps2, npin = oracle_lib.pin_boundary(v, t, phi, psi)
    p = oracle_lib.make_params(Phys, verts=v, tris=t, mirror=True, wall_guards=True, n_pin=npin)
    sim = oracle_lib.Sim(p, v2, t2, np.zeros(len(v2)) if ph2 is None else ph2)
    if ps2 is not None:
        sim.psi = ps2.copy()
    return sim, p, npin


def test_pinned_compute_psi_hemisphere(ctx, oracle_lib):
    say("pinned compute_psi hemisphere nu=8")
    v, t = meshgen.hemisphere(8)
    rng = np.random.default_rng(0)
    phi, psi = 1.0 + 0.1 * rng.normal(size=len(v)), -1.0 + 0.1 * rng.normal(size=len(v))
    sim = _gpu(ctx, v, t, "lu64", phi, psi)
    osim, p, npin = _oracle(oracle_lib, v, t, phi, psi)
    assert sim.N_pin == npin == 32
    x, ph, ps, _ = sim.state()
    assert np.array_equal(x, osim.x) and np.array_equal(ph, osim.phi) and np.array_equal(ps, osim.psi)
    assert np.array_equal(sim.get_triangles(), osim.t)
    got = sim.compute_psi()
    want = oracle_lib.compute_psi(p, osim.x, osim.t, osim.phi)
    assert rel(got, want) < 1e-10


@pytest.mark.parametrize("solver,tol", [("lu64", 1e-10), ("lu16x3_ir64", 1e-8), ("lu16_ir64", 1e-8)])
def test_pinned_rk4_reference_input(ctx, oracle_lib, solver, tol):
    """two RK4 steps (dt = 1e-3, fixed) of the reference's pinned bubble from
    rest (phi = 0, as pinned-bubble.cpp:123), wall mode, vs the oracle"""
    say(f"pinned rk4 init-pinned.ply {solver}")
    v, t, _, _ = bemhip.read_ply(PINNED_PLY)
    sim = _gpu(ctx, v, t, solver)
    osim, _, npin = _oracle(oracle_lib, v, t)
    assert sim.N_pin == npin == 80
    x0, _, _, _ = sim.state()
    assert np.array_equal(x0, osim.x)
    for _ in range(2):
        sim.evolve_system_RK4(1e-3, fixdt=True)
        osim.rk4(1e-3, True)
    x, ph, ps, tm = sim.state()
    assert abs(tm - osim.time) < 1e-15
    assert np.all(x[-npin:, 0] == 0.0)                          # pinned vertices stay on the wall
    assert rel(x - x0, osim.x - x0) < tol                       # displacements, pinned drift included
    assert rel((x - x0)[-npin:], (osim.x - x0)[-npin:]) < tol
    assert rel(ph, osim.phi) < tol and rel(ps, osim.psi) < tol


def test_pinned_euler_and_position_t_hemisphere(ctx, oracle_lib):
    say("pinned euler + position_t hemisphere nu=6")
    v, t = meshgen.hemisphere(6)
    rng = np.random.default_rng(2)
    phi, psi = 0.1 * rng.normal(size=len(v)), 0.1 * rng.normal(size=len(v))
    sim = _gpu(ctx, v, t, "lu64", phi, psi)
    osim, p, npin = _oracle(oracle_lib, v, t, phi, psi)
    vel, psil = sim.position_t()
    x0, ph0, _, _ = sim.state()
    assert np.array_equal(ph0, osim.phi)                  # position_t works on a copy of phi
    sim.evolve_system(1e-3, fixdt=True)
    osim.euler(1e-3, True)
    x, ph, _, _ = sim.state()
    assert rel(x, osim.x) < 1e-10 and rel(ph, osim.phi) < 1e-10
    assert np.max(np.abs((x - x0) / 1e-3 - vel)) < 1e-9 * np.max(np.abs(vel))
    assert np.all(vel[-npin:, 0] == 0.0)


def test_export_and_restart(ctx, tmp_path):
    """export_mesh (float, the reference's default) and the lossless double
    file; a ColocSimPin restarted from the double file (pinned-bubble-restart
    .cpp:66-124) continues on the same state up to the reordering of the
    pinned block."""
    say("pinned export + restart")
    v, t = meshgen.hemisphere(6)
    rng = np.random.default_rng(4)
    sim = _gpu(ctx, v, t, "lu64", 0.05 * rng.normal(size=len(v)))
    V0 = sim.get_volume()
    sim.evolve_system_RK4(1e-3, fixdt=True)
    x, ph, ps, _ = sim.state()
    tri = sim.get_triangles()
    f8, f4 = str(tmp_path / "s8.ply"), str(tmp_path / "s4.ply")
    sim.export_mesh(f8, precision=8)
    sim.export_mesh(f4)
    v8, t8, p8, s8 = bemhip.read_ply(f8)
    assert np.array_equal(v8, x) and np.array_equal(t8, tri) and np.array_equal(p8, ph) and np.array_equal(s8, ps)
    v4, _, p4, _ = bemhip.read_ply(f4)
    assert np.array_equal(v4, x.astype(np.float32).astype(np.float64))
    assert np.array_equal(p4, ph.astype(np.float32).astype(np.float64))
    # restart; V_0 would default to the restart mesh's volume (Simulation.hpp:42),
    # so it is carried over, as a driver does with set_V_0
    r = _gpu(ctx, v8, t8, "lu64", p8, s8, V_0=V0)
    r.set_time(sim.get_time())
    rv, rt, rp, rs, npin = bemhip.pin_boundary(v8, t8, p8, s8)
    xr, phr, psr, _ = r.state()
    assert r.N_pin == npin == sim.N_pin
    assert np.array_equal(xr, rv) and np.array_equal(phr, rp) and np.array_equal(psr, rs)
    # same vertices, the pinned block permuted: map restart ids back to the original ones
    key = {tuple(p): i for i, p in enumerate(x)}
    perm = np.array([key[tuple(p)] for p in xr])
    assert np.array_equal(perm[:-npin], np.arange(len(x) - npin))
    sim.evolve_system_RK4(1e-3, fixdt=True)
    r.evolve_system_RK4(1e-3, fixdt=True)
    xa, pa, sa, _ = sim.state()
    xb, pb, sb, _ = r.state()
    assert rel(xb, xa[perm]) < 1e-10 and rel(pb, pa[perm]) < 1e-10 and rel(sb, sa[perm]) < 1e-10


def test_pinned_exterior_pot_and_guards(ctx, oracle_lib):
    say("pinned exterior_pot + guards")
    v, t = meshgen.hemisphere(6)
    sim = _gpu(ctx, v, t, "lu64", np.ones(len(v)), -np.ones(len(v)))
    osim, p, npin = _oracle(oracle_lib, v, t, np.ones(len(v)), -np.ones(len(v)))
    pts = np.array([[1.5, 0.2, 0.1], [2.0, -0.5, 0.3], [0.3, 0.1, 0.0]])
    got = sim.exterior_pot(pts)
    # LinLinSim::exterior_pot on a ColocSimPin: the swapped solve's vector goes in as psi
    xs = oracle_lib.compute_psi(p, osim.x, osim.t, osim.phi)
    want = oracle_lib.exterior_pot(osim.x, osim.t, osim.phi, xs, pts)
    assert rel(got, want) < 1e-10
    with pytest.raises(bemhip.BemHipError):
        sim.adjoint_radius(2, 1e-3, np.ones(2), every=1)
    with pytest.raises(bemhip.BemHipError):
        sim.N_pin = len(v) + 1
