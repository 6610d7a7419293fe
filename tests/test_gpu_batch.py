"""Batches of independent bubbles (BEMHIP_INDEPENDENT_PARTS): every loose part
is its own BEM problem with its own V_0 and forcing, advanced with a shared dt.

Parity: the batch must reproduce, part by part, the oracle run of each bubble
as a separate simulation (FP64 path 1e-10 relative in psi and phi, 1e-12 in
x; mixed path 1e-8 in psi), and the batched compute_psi must equal the
single-part compute_psi of the same library (1e-12: only the summation order
of the assembly differs).
"""

import os
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


def _cases(seed=0):
    """Three uneven spheres (N = 162, 642, 362) with SH-perturbed potentials
    and their own forcing amplitudes."""
    out = []
    for k, (nu, R, A) in enumerate([(4, 0.7, 0.2), (8, 1.3, 0.45), (6, 1.0, 0.1)]):
        v, t = meshgen.icosphere(nu, R)
        v = v + 1e-3 * np.random.default_rng(seed + k).standard_normal(v.shape)
        phi = -0.3 + 0.05 * meshgen.random_sh_field(v, 1, 3, seed + 10 + k)
        out.append(meshgen.Case(f"b{k}", np.ascontiguousarray(v), t, phi,
                                meshgen.Physics(1.0, 1.0, 0.0, 1.0, field_amp=A, field_freq=0.5), dt=1e-2))
    return out


@pytest.mark.parametrize("solver,tol", [("lu64", 1e-12), ("lu16x3_ir64", 1e-8), ("lu16_ir64", 1e-8)])
def test_batched_compute_psi_equals_single(ctx, oracle_lib, solver, tol):
    say(f"batch compute_psi {solver}")
    cases = _cases()
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    psi, st = ctx.compute_psi(v, t, phi, solver, independent_parts=True)
    assert st["converged"]
    for c, (o, n) in zip(cases, parts):
        one, _ = ctx.compute_psi(c.verts, c.tris, c.phi, solver)
        assert rel(psi[o:o + n], one) < tol
        p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
        ref = oracle_lib.compute_psi(p, c.verts, c.tris, c.phi)
        assert rel(psi[o:o + n], ref) < (1e-10 if solver == "lu64" else 1e-8)
    # coupled (the reference's semantics for several parts) differs: the parts interact
    coupled, _ = ctx.compute_psi(v, t, phi, solver)
    assert rel(coupled, psi) > 1e-6


@pytest.mark.parametrize("solver", ["lu64", "lu16x3_ir64", "lu16_ir64"])
def test_batched_rk4_equals_separate_simulations(ctx, oracle_lib, solver):
    say(f"batch rk4 {solver}")
    cases = _cases(seed=3)
    v, t, phi, parts, amp, freq = meshgen.join_batch(cases)
    s = bemhip.ColocSim(ctx, v, t, 1.0, 1.0, 0.0, 1.0, solver=solver, independent_parts=True)
    s.set_phi(phi)
    assert s.nparts() == len(cases)
    vols = s.part_volumes()
    s.set_part_physics(field_amp=amp, field_freq=freq)   # V_0 stays per part (own initial volume)
    s.set_time(0.3)                                       # forcing is nonzero from the first step
    for _ in range(2):
        s.evolve_system_RK4(1e-2, True)
    x, ph, psi, tm = s.state()
    for c, (o, n), V in zip(cases, parts, vols):
        p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
        ref = oracle_lib.Sim(p, c.verts, c.tris, c.phi)
        assert abs(V - oracle_lib.part_volumes(c.verts, c.tris).sum()) < 1e-12 * V
        ref.time = 0.3
        for _ in range(2):
            ref.rk4(1e-2, True)
        off = v[o:o + n] - c.verts          # grid placement of this part in the batch
        xb = x[o:o + n] - off
        if solver == "lu64":
            assert rel(xb, ref.x) < 1e-12 and rel(ph[o:o + n], ref.phi) < 1e-10 and rel(psi[o:o + n], ref.psi) < 1e-10
        else:
            assert rel(xb, ref.x) < 1e-10 and rel(ph[o:o + n], ref.phi) < 1e-8 and rel(psi[o:o + n], ref.psi) < 1e-8
    assert abs(tm - 0.32) < 1e-15


@pytest.fixture(scope="module")
def ctx_lanes():
    """A context whose batches solve every problem on the solver lanes
    (the one-launch small-problem solver switched off)."""
    old = os.environ.get("BEMHIP_SMALL_N")
    os.environ["BEMHIP_SMALL_N"] = "0"
    try:
        c = bemhip.Context(0)
    finally:
        if old is None:
            del os.environ["BEMHIP_SMALL_N"]
        else:
            os.environ["BEMHIP_SMALL_N"] = old
    yield c
    c.close()


@pytest.mark.parametrize("solver", ["lu64", "lu16x3_ir64"])
def test_small_problem_solver_matches_lanes_and_oracle(ctx, ctx_lanes, oracle_lib, solver):
    """Bubbles of at most 1024 unknowns are solved in one launch (FP64 GEPP
    and triangular solves, one work-group per bubble); a bubble of N = 1442
    stays on the solver lanes in the same batch.  Both paths agree part by
    part, and the small bubbles match the oracle's FP64 LU for either
    solver (the small path has no FP32 factor)."""
    say(f"batch small-problem solver {solver}")
    cases = _cases(seed=5)
    v12, t12 = meshgen.icosphere(12, 1.1)
    v12 = v12 + 1e-3 * np.random.default_rng(9).standard_normal(v12.shape)
    cases.append(meshgen.Case("b12", np.ascontiguousarray(v12), t12,
                              -0.2 + 0.05 * meshgen.random_sh_field(v12, 1, 3, 19),
                              meshgen.Physics(1.0, 1.0, 0.0, 1.0), dt=1e-2))
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    psi, st = ctx.compute_psi(v, t, phi, solver, independent_parts=True)
    psi_l, st_l = ctx_lanes.compute_psi(v, t, phi, solver, independent_parts=True)
    assert st["converged"] and st_l["converged"]
    for c, (o, n) in zip(cases, parts):
        small = n <= 1024
        assert rel(psi[o:o + n], psi_l[o:o + n]) < (1e-12 if solver == "lu64" else 1e-8)
        p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
        ref = oracle_lib.compute_psi(p, c.verts, c.tris, c.phi)
        assert rel(psi[o:o + n], ref) < (1e-10 if (small or solver == "lu64") else 1e-8)


@pytest.fixture(scope="module")
def ctx_small1():
    """A context whose single-problem LU solves of at most 1024 unknowns run
    on the one-work-group solver of the batches (BEMHIP_SMALL_SINGLE=1)."""
    old = os.environ.get("BEMHIP_SMALL_SINGLE")
    os.environ["BEMHIP_SMALL_SINGLE"] = "1"
    try:
        c = bemhip.Context(0)
    finally:
        if old is None:
            del os.environ["BEMHIP_SMALL_SINGLE"]
        else:
            os.environ["BEMHIP_SMALL_SINGLE"] = old
    yield c
    c.close()


@pytest.mark.parametrize("n", [13, 300, 1024])
def test_small_solver_row_interchanges(ctx_small1, n):
    """The one-work-group solver on matrices that need row interchanges
    (random normal entries; BEM matrices never pivot), against LAPACK, for
    both an FP64 and a mixed solver id (the small path is FP64 for both)."""
    say(f"small solver interchanges n={n}")
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    ref = np.linalg.solve(A, b)
    for solver in ("lu64", "lu16x3_ir64"):
        x, st = ctx_small1.solve(A, b, solver)
        assert rel(x, ref) < 1e-10 and st["panels"] == 0


def test_small_solver_zero_pivot(ctx_small1):
    say("small solver zero pivot")
    A = np.eye(20)
    A[:, 7] = 0.0   # a zero column: GEPP finds a zero pivot there
    with pytest.raises(bemhip.BemHipError) as e:
        ctx_small1.solve(A, np.ones(20), "lu64")
    assert e.value.code == bemhip.ERR_SINGULAR and "column 8" in str(e.value)


def test_batch_part_physics_reset_and_errors(ctx):
    say("batch part physics")
    cases = _cases()
    v, t, phi, parts, amp, freq = meshgen.join_batch(cases)
    s = bemhip.ColocSim(ctx, v, t, solver="lu64", independent_parts=True)
    s.set_phi(phi)
    with pytest.raises(bemhip.BemHipError):
        s.ctx.check(s.ctx.L.bemhip_sim_set_part_physics(s.h, 2, None, None, None, None))   # wrong part count
    # per-part V_0 equal to the scalar reproduces the scalar run
    vel = np.zeros((s.N, 3))
    s.set_V_0(5.0)
    d_scalar = s.pot_t_multi(vel, 0.0)
    s.set_part_physics(V_0=np.full(3, 5.0))
    assert rel(s.pot_t_multi(vel, 0.0), d_scalar) == 0.0
    # the coupled mode cannot be switched on after creation
    p = bemhip.SimParams()
    s.ctx.L.bemhip_sim_default_params(p)
    p.solver = 0
    with pytest.raises(bemhip.BemHipError):
        s.ctx.check(s.ctx.L.bemhip_sim_set_params(s.h, p))


def test_batch_rejects_interleaved_parts(ctx):
    say("batch interleaved")
    cases = _cases()
    v, t, phi, parts, _, _ = meshgen.join_batch(cases[:2])
    # move the last vertex of part 0 behind part 1
    n0 = parts[0][1]
    order = np.r_[0:n0 - 1, n0:v.shape[0], n0 - 1]
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    v2, t2, phi2 = v[order], inv[t].astype(np.uint32), phi[order]
    with pytest.raises(bemhip.BemHipError) as e:
        ctx.compute_psi(v2, t2, phi2, "lu64", independent_parts=True)
    assert e.value.code == 2   # BEMHIP_ERR_BAD_MESH
    psi, _ = ctx.compute_psi(v2, t2, phi2, "lu64")   # the coupled problem accepts any order
    assert np.all(np.isfinite(psi))
