"""Potential at points off the surface on the GPU (compute_exterior_pot,
LinLinSim.cpp:440-477; LinLinSim::exterior_pot, LinLinSim.cpp:478-483)
against the FP64 oracle, plus Green's identities at a size the oracle would
take minutes for."""

import sys

import numpy as np
import pytest

import bemhip
import meshgen

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _points(rng, count, center, lo, hi):
    """count points at distances in [lo, hi] from center, none within 0.05 of the unit sphere there"""
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.asarray(center) + d * rng.uniform(lo, hi, size=(count, 1))


def test_exterior_potential_matches_oracle(ctx, oracle_lib):
    say("exterior potential nu=16 P=611")
    c = (0.1, -0.2, 0.3)
    v, t = meshgen.icosphere(16, 1.0, c)
    rng = np.random.default_rng(11)
    phi, psi = rng.normal(size=len(v)), rng.normal(size=len(v))
    # a ragged count (not a multiple of the 256-point work-group), inside and out
    pts = np.concatenate([_points(rng, 400, c, 1.05, 4.0), _points(rng, 211, c, 0.0, 0.95)])
    got = ctx.exterior_potential(v, t, phi, psi, pts)
    want = oracle_lib.exterior_pot(v, t, phi, psi, pts)
    assert rel(got, want) < 1e-12
    # a single point, and none
    one = ctx.exterior_potential(v, t, phi, psi, pts[:1])
    assert abs(one[0] - want[0]) <= 1e-12 * max(1.0, abs(want[0]))
    assert ctx.exterior_potential(v, t, phi, psi, np.zeros((0, 3))).shape == (0,)


def test_exterior_potential_green_identities_large(ctx):
    """nu = 32 sphere (N = 10242), 4099 points: phi = 1, psi = -1 gives 1/|y|
    outside and 0 inside to O(h^2); phi = 1, psi = 0 gives 0 outside and -1
    inside to the rule's error only."""
    say("exterior potential Green identities nu=32 P=4099")
    v, t = meshgen.icosphere(32, 1.0)
    rng = np.random.default_rng(3)
    out_pts, in_pts = _points(rng, 3000, (0, 0, 0), 1.5, 3.0), _points(rng, 1099, (0, 0, 0), 0.0, 0.5)
    pts = np.concatenate([out_pts, in_pts])
    one = np.ones(len(v))
    f = ctx.exterior_potential(v, t, one, -one, pts)
    assert np.max(np.abs(f[:3000] - 1.0 / np.linalg.norm(out_pts, axis=1))) < 3e-4   # nu=8 gives 2.8e-3
    assert np.max(np.abs(f[3000:])) < 3e-4
    cst = ctx.exterior_potential(v, t, one, 0 * one, pts)
    assert np.max(np.abs(cst[:3000])) < 1e-12 and np.max(np.abs(cst[3000:] + 1.0)) < 1e-12


def test_exterior_potential_bad_inputs(ctx):
    v, t = meshgen.icosphere(2, 1.0)
    one = np.ones(len(v))
    bad = t.copy()
    bad[0, 0] = len(v)
    with pytest.raises(bemhip.BemHipError):
        ctx.exterior_potential(v, bad, one, one, np.array([[2.0, 0, 0]]))
    with pytest.raises(bemhip.BemHipError):
        ctx.exterior_potential(v, t, one, one, np.array([[np.nan, 0, 0]]))


@pytest.mark.parametrize("solver,tol", [("lu64", 1e-10), ("lu16x3_ir64", 1e-8), ("lu16_ir64", 1e-8)])
def test_sim_exterior_pot_matches_oracle(ctx, oracle_lib, solver, tol):
    say(f"sim exterior_pot {solver}")
    case = meshgen.config_small(seed=0)
    ph = case.physics
    s = bemhip.ColocSim(ctx, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, solver=solver)
    s.set_phi(case.phi)
    psi_before = s.get_psi()
    rng = np.random.default_rng(4)
    pts = np.concatenate([_points(rng, 300, (0, 0, 0), 1.1, 5.0), _points(rng, 33, (0, 0, 0), 0.0, 0.9)])
    got = s.exterior_pot(pts)
    assert np.array_equal(s.get_psi(), psi_before)            # the simulation's psi is left as it was
    p = oracle_lib.make_params(ph, verts=case.verts, tris=case.tris)
    psi = oracle_lib.compute_psi(p, case.verts, case.tris, case.phi)
    want = oracle_lib.exterior_pot(case.verts, case.tris, case.phi, psi, pts)
    assert rel(got, want) < tol
    # the same as the standalone entry on the simulation's own psi
    s.compute_psi()
    alone = ctx.exterior_potential(case.verts, case.tris, case.phi, s.get_psi(), pts)
    assert rel(got, alone) < 1e-13
