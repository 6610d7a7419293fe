"""GPU parity of the gfx950 path against the FP64 CPU oracle (calls the C-ABI).

Tolerances (written per test): assembly entries 1e-12 relative to the
largest entry of the matrix, plain FP64 solves / stages 1e-10 relative, mixed
precision (FP32 or split-fp16 LU + FP64 refinement) 1e-8 relative in psi.
"""

import math
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


@pytest.mark.parametrize("mirror", [False, True])
@pytest.mark.parametrize("nu", [3, 8])
def test_assembly_matches_oracle(ctx, oracle_lib, nu, mirror):
    say(f"assembly nu={nu} mirror={mirror}")
    center = (1.6, 0.1, -0.2) if mirror else (0.1, 0.2, -0.3)
    v, t = meshgen.icosphere(nu, 1.1, center)
    v = v + 1e-3 * np.random.default_rng(nu).standard_normal(v.shape)
    phi = 1.0 + 0.1 * meshgen.random_sh_field(v, 1, 3, 2, center)
    G, H, hphi = ctx.assemble(v, t, mirror=mirror, want_H=True, phi=phi)
    Go, Ho = oracle_lib.assemble(v, t, mirror=mirror)
    assert rel(G, Go) < 1e-12
    assert rel(H, Ho) < 1e-12
    assert rel(hphi, Ho @ phi) < 1e-12
    if not mirror:
        assert np.max(np.abs(H.sum(1) + 4 * math.pi)) < 1e-11


def test_lu64_with_pivoting(ctx):
    say("lu64 random N=1000")
    rng = np.random.default_rng(5)
    A = rng.standard_normal((1000, 1000))
    b = rng.standard_normal(1000)
    x, st = ctx.solve(A, b, "lu64")
    assert rel(x, np.linalg.solve(A, b)) < 1e-10
    # the speculative (interchange-free) panel is rejected and GEPP takes over
    assert st["panels"] >= 2 and st["panels_pivoted"] == st["panels"]
    for solver in ("lu32_ir64", "lu16x3_ir64"):
        xm, stm = ctx.solve(A, b, solver)
        assert rel(xm, np.linalg.solve(A, b)) < 1e-10 and stm["panels_pivoted"] >= 1


def test_lu_accepted_then_pivoted_panels(ctx):
    """Columns [0, 1536) diagonally dominant, the rest random: the first panels
    are accepted without interchanges and later ones fall back to partial
    pivoting, whose row swaps reach into L21 blocks that the second stream
    copied into the factor.  The fp16 LU's merged schedule checks its panels'
    verdicts once at the end and redoes the factorisation step by step, where
    the same per-panel fallback applies (the discarded pass is not counted)."""
    say("lu accepted then pivoted N=3000")
    rng = np.random.default_rng(11)
    n = 3000
    A = rng.standard_normal((n, n))
    A[np.arange(1536), np.arange(1536)] += 100.0
    b = rng.standard_normal(n)
    ref = np.linalg.solve(A, b)
    for solver in ("lu64", "lu32_ir64", "lu16x3_ir64", "lu16_ir64"):
        x, st = ctx.solve(A, b, solver)
        say(f"  {solver}: panels {st['panels']} pivoted {st['panels_pivoted']}")
        assert rel(x, ref) < 1e-10
        assert 0 < st["panels_pivoted"] < st["panels"]


def test_speculative_panels_equal_gepp(oracle_lib, monkeypatch):
    """On BEM matrices partial pivoting keeps the diagonal: the interchange-free
    panels are accepted and give GEPP's factorisation (checked against a
    context that always runs the GEPP panels)."""
    say("speculation vs gepp")
    v, t = meshgen.icosphere(12)
    G, H = oracle_lib.assemble(v, t)
    b = H @ (1.0 + 0.05 * meshgen.random_sh_field(v, 1, 4, 5))
    c1 = bemhip.Context(0)
    x1, st1 = c1.solve(G, b, "lu64")
    c1.close()
    monkeypatch.setenv("BEMHIP_LU_SPECULATE", "0")
    c2 = bemhip.Context(0)
    x2, st2 = c2.solve(G, b, "lu64")
    c2.close()
    assert st1["panels_pivoted"] == 0 and st2["panels_pivoted"] == st2["panels"] > 0
    assert rel(x1, x2) < 1e-12
    assert rel(x1, oracle_lib.solve(G, b)) < 1e-10
    # the top blocks by 32-column step launches instead of one persistent launch
    monkeypatch.setenv("BEMHIP_LU_SPECULATE", "1")
    monkeypatch.setenv("BEMHIP_NP_TOP", "0")
    c3 = bemhip.Context(0)
    for solver, tol in (("lu64", 1e-12), ("lu16x3_ir64", 1e-10), ("lu16_ir64", 1e-8)):
        x3, st3 = c3.solve(G, b, solver)
        assert st3["panels_pivoted"] == 0 and rel(x3, x1) < tol
    c3.close()


def test_split_fixed_scale_and_redo(oracle_lib, monkeypatch):
    """The split-fp16 updates use one B scale per factorisation (taken from
    max |diag G|) while every U12 block fits its window.  On a BEM matrix the
    result agrees with per-block scales to the refinement tolerance; on a matrix
    whose U12 blocks are ~1000x its diagonal the check fails and the
    factorisation is redone with per-block scales, which then gives exactly
    the per-block result."""
    say("split-fp16 fixed scale vs per-block scales")
    v, t = meshgen.icosphere(12)
    G, H = oracle_lib.assemble(v, t)
    b = H @ (1.0 + 0.05 * meshgen.random_sh_field(v, 1, 4, 5))
    rng = np.random.default_rng(8)
    A = rng.normal(size=(2600, 2600))
    A[np.diag_indices(2600)] *= 1e-3
    bA = rng.normal(size=2600)
    c1 = bemhip.Context(0)
    xg1, _ = c1.solve(G, b, "lu16x3_ir64")
    xa1, sa1 = c1.solve(A, bA, "lu16x3_ir64")
    c1.close()
    monkeypatch.setenv("BEMHIP_SPLIT_FIXED", "0")
    c2 = bemhip.Context(0)
    xg2, _ = c2.solve(G, b, "lu16x3_ir64")
    xa2, sa2 = c2.solve(A, bA, "lu16x3_ir64")
    c2.close()
    ref = oracle_lib.solve(G, b)
    assert rel(xg1, ref) < 1e-8 and rel(xg2, ref) < 1e-8
    assert np.array_equal(xa1, xa2) and sa1["iterations"] == sa2["iterations"]
    assert rel(xa1, np.linalg.solve(A, bA)) < 1e-8


def test_lu16_with_pivoting(ctx):
    """fp16 products with GEPP panels: a row-permuted diagonally dominant
    matrix (well conditioned, but partial pivoting must interchange rows in
    every panel) is solved to the FP64 refinement's accuracy."""
    say("lu16 pivoted panels N=2600")
    rng = np.random.default_rng(21)
    n = 2600
    D = rng.standard_normal((n, n)) + 3.0 * np.sqrt(n) * np.eye(n)
    A = D[rng.permutation(n)]
    b = rng.standard_normal(n)
    x, st = ctx.solve(A, b, "lu16_ir64")
    say(f"  {st}")
    assert st["converged"] == 1 and st["panels_pivoted"] == st["panels"] > 1
    assert rel(x, np.linalg.solve(A, b)) < 1e-10


def test_lu16_overflow_falls_back_to_fp32(ctx):
    """Operands scaled from max |diag A| overflow fp16 when the off-diagonal
    entries are ~1000x the diagonal: the factorisation is redone with FP32
    products, which then is the lu32_ir64 factorisation exactly."""
    say("lu16 fp16 overflow -> FP32 redo")
    rng = np.random.default_rng(8)
    A = rng.normal(size=(2600, 2600))
    A[np.diag_indices(2600)] *= 1e-3
    b = rng.normal(size=2600)
    x16, s16 = ctx.solve(A, b, "lu16_ir64")
    x32, s32 = ctx.solve(A, b, "lu32_ir64")
    assert np.array_equal(x16, x32) and s16["iterations"] == s32["iterations"]
    assert rel(x16, np.linalg.solve(A, b)) < 1e-8


@pytest.mark.parametrize("solver,tol", [("lu64", 1e-10), ("lu32_ir64", 1e-8), ("lu16x3_ir64", 1e-8),
                                        ("lu16_ir64", 1e-8), ("bicgstab64", 1e-10)])
def test_solvers_on_assembled_g(ctx, oracle_lib, solver, tol):
    say(f"solve {solver}")
    v, t = meshgen.icosphere(12)
    G, H = oracle_lib.assemble(v, t)
    b = H @ (1.0 + 0.05 * meshgen.random_sh_field(v, 1, 4, 3))
    x, st = ctx.solve(G, b, solver)
    ref = oracle_lib.solve(G, b)
    assert rel(x, ref) < tol, st
    if solver.endswith("ir64"):
        assert st["converged"] == 1 and st["iterations"] >= 1


@pytest.mark.parametrize("solver,tol", [("lu64", 1e-10), ("lu32_ir64", 1e-8), ("bicgstab64", 1e-10)])
def test_compute_psi_matches_oracle(ctx, oracle_lib, solver, tol):
    say(f"compute_psi {solver}")
    c = meshgen.config_small(seed=0)
    p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
    ref = oracle_lib.compute_psi(p, c.verts, c.tris, c.phi)
    psi, st = ctx.compute_psi(c.verts, c.tris, c.phi, solver)
    assert rel(psi, ref) < tol, st


def _oracle_run(oracle_lib, case, steps, solver="lu", mirror=False, frozen=True, rk4=True):
    p = oracle_lib.make_params(case.physics, verts=case.verts, tris=case.tris, solver=solver, mirror=mirror,
                               wall_guards=mirror, frozen_stage_time=frozen)
    sim = oracle_lib.Sim(p, case.verts, case.tris, case.phi)
    for _ in range(steps):
        (sim.rk4 if rk4 else sim.euler)(case.dt, True)
    return sim


def _gpu_sim(ctx, case, solver, mirror=False, frozen=True):
    ph = case.physics
    s = bemhip.ColocSim(ctx, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma,
                        field_amp=ph.field_amp, field_freq=ph.field_freq, mirror=mirror,
                        frozen_stage_time=frozen, solver=solver)
    s.set_phi(case.phi)
    return s


def test_rk4_step_config_small(ctx, oracle_lib):
    """One RK4 step of the nu=8 sphere (dt = 1e-3, LU), FP64 path vs oracle."""
    say("rk4 small")
    c = meshgen.config_small(seed=0)
    ref = _oracle_run(oracle_lib, c, 1)
    s = _gpu_sim(ctx, c, "lu64")
    s.evolve_system_RK4(c.dt, True)
    x, phi, psi, t = s.state()
    assert rel(x, ref.x) < 1e-12
    assert rel(phi, ref.phi) < 1e-10
    assert rel(psi, ref.psi) < 1e-10
    assert abs(t - ref.time) < 1e-15


def test_rk4_mixed_precision_config_small(ctx, oracle_lib):
    say("rk4 small mixed")
    c = meshgen.config_small(seed=1)
    ref = _oracle_run(oracle_lib, c, 1)
    s = _gpu_sim(ctx, c, "lu32_ir64")
    s.evolve_system_RK4(c.dt, True)
    x, phi, psi, t = s.state()
    assert rel(psi, ref.psi) < 1e-8 and rel(phi, ref.phi) < 1e-8 and rel(x, ref.x) < 1e-10


def test_two_bubbles_forced_classical_stages(ctx, oracle_lib):
    """Two loose parts (per-part volumes), time-dependent forcing, stage times
    advanced (frozen_stage_time off) and adaptive dt: 2 RK4 steps."""
    say("two bubbles")
    c = meshgen.config_two_bubbles(seed=0, nu=4)
    c.physics.field_amp, c.physics.field_freq = 0.3, 0.5
    c.dt = 0.05
    p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris, frozen_stage_time=False)
    ref = oracle_lib.Sim(p, c.verts, c.tris, c.phi)
    s = _gpu_sim(ctx, c, "lu64", frozen=False)
    for _ in range(2):
        dt_ref = ref.rk4(c.dt, False)
        dt = s.evolve_system_RK4(c.dt, False)
        assert abs(dt - dt_ref) < 1e-12 * dt_ref
    x, phi, psi, t = s.state()
    assert rel(x, ref.x) < 1e-11 and rel(phi, ref.phi) < 1e-10 and rel(psi, ref.psi) < 1e-10


def test_wall_mode_rk4_and_negative_x(ctx, oracle_lib):
    say("wall mode")
    v, t = meshgen.icosphere(4, 0.5, (0.8, 0.0, 0.0))
    phi = np.full(len(v), -0.05)
    case = meshgen.Case("wall", v, t, phi, meshgen.Physics(), dt=1e-2)
    ref = _oracle_run(oracle_lib, case, 2, mirror=True)
    s = _gpu_sim(ctx, case, "lu64", mirror=True)
    for _ in range(2):
        s.evolve_system_RK4(case.dt, True)
    x, phi_g, psi, _ = s.state()
    assert rel(x, ref.x) < 1e-11 and rel(phi_g, ref.phi) < 1e-10 and rel(psi, ref.psi) < 1e-10
    bad = v.copy()
    bad[0, 0] = -0.01
    s2 = bemhip.ColocSim(ctx, bad, t, mirror=True, solver="lu64")
    with pytest.raises(IndexError):
        s2.evolve_system_RK4(1e-3, True)


def test_euler_and_bicgstab_default(ctx, oracle_lib):
    say("euler")
    c = meshgen.config_small(seed=2)
    ref = _oracle_run(oracle_lib, c, 2, solver="bicgstab", rk4=False)
    s = _gpu_sim(ctx, c, "bicgstab64")
    for _ in range(2):
        s.evolve_system(c.dt, True)
    x, phi, _, t = s.state()
    assert rel(x, ref.x) < 1e-11 and rel(phi, ref.phi) < 1e-10


def test_surface_tension_curvature_path(ctx, oracle_lib):
    say("sigma != 0")
    c = meshgen.config_small(seed=3)
    c.physics.sigma = 0.7
    vel = np.random.default_rng(0).standard_normal((len(c.verts), 3))
    p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
    ref = oracle_lib.pot_t_multi(p, c.verts, c.tris, vel, 0.25)
    s = _gpu_sim(ctx, c, "lu64")
    got = s.pot_t_multi(vel, 0.25)
    assert rel(got, ref) < 1e-12
    # arrays of the wrong length are refused before the library reads them
    with pytest.raises(ValueError):
        s.pot_t_multi(vel[:-1], 0.25)
    with pytest.raises(ValueError):
        s.set_psi(np.zeros(len(c.verts) - 1))


def test_bad_inputs_are_reported(ctx):
    say("bad inputs")
    v, t = meshgen.icosphere(2)
    bad = t.copy()
    bad[0, 0] = 10 ** 6
    with pytest.raises(bemhip.BemHipError) as e:
        ctx.compute_psi(v, bad, np.ones(len(v)))
    assert e.value.code == bemhip.ERR_BAD_MESH
    phi = np.ones(len(v))
    phi[3] = np.nan
    with pytest.raises(bemhip.BemHipError) as e:
        ctx.compute_psi(v, t, phi)
    assert e.value.code == bemhip.ERR_NAN


def test_triangular_solve_give_up_is_reported(oracle_lib, monkeypatch):
    """The persistent triangular solves wait on each other's blocks with a
    bounded spin; a work-group that gives up finishes with invalid results and
    raises a sticky error word.  Here the host pre-sets the error word before
    every launch (BEMHIP_TRSV_FORCE_GIVEUP=1), so every waiting work-group gives
    up at once and exits normally: each LU solver must then report
    BEMHIP_ERR_HIP instead of returning x.  A context without the hook solves
    the same system afterwards."""
    say("triangular solve give-up")
    v, t = meshgen.icosphere(12)   # N = 1442: 12 row blocks of 128
    G, H = oracle_lib.assemble(v, t)
    b = H @ (1.0 + 0.05 * meshgen.random_sh_field(v, 1, 4, 5))
    monkeypatch.setenv("BEMHIP_TRSV_FORCE_GIVEUP", "1")
    c1 = bemhip.Context(0)
    try:
        for solver in ("lu64", "lu32_ir64", "lu16_ir64"):
            with pytest.raises(bemhip.BemHipError) as e:
                c1.solve(G, b, solver)
            assert e.value.code == bemhip.ERR_HIP, solver
            assert "gave up" in str(e.value)
    finally:
        c1.close()
    monkeypatch.delenv("BEMHIP_TRSV_FORCE_GIVEUP")
    c2 = bemhip.Context(0)
    try:
        x, st = c2.solve(G, b, "lu64")
        assert rel(x, oracle_lib.solve(G, b)) < 1e-10
    finally:
        c2.close()


def test_large_sphere_properties(ctx):
    """nu = 32 (N = 10242): analytic sphere solution and agreement of the
    mixed-precision paths with the FP64 LU at full size (no oracle)."""
    say("large nu=32")
    v, t = meshgen.icosphere(32)
    psi64, st64 = ctx.compute_psi(v, t, np.ones(len(v)), "lu64")
    # psi = -phi0 / R up to the O(h^2) discretisation error (4.2e-3 at nu = 8, /16 at nu = 32)
    assert abs(psi64.mean() + 1.0) < 4e-4 and np.std(psi64) < 1e-4
    Y = meshgen.real_sph_harm(2, v)[(2, 1)]
    phi = -1.0 + 0.02 * Y
    p64, _ = ctx.compute_psi(v, t, phi, "lu64")
    for solver in ("lu32_ir64", "lu16x3_ir64", "lu16_ir64"):
        p, st = ctx.compute_psi(v, t, phi, solver)
        say(solver, st)
        assert rel(p, p64) < 1e-8
    # linearity: psi(phi) = psi(-1) + 0.02 psi(Y)
    pY, _ = ctx.compute_psi(v, t, Y, "lu64")
    assert rel(p64, -psi64 + 0.02 * pY) < 1e-11


def test_large_speculative_lu_equals_gepp(monkeypatch):
    """nu = 32: the look-ahead LU with interchange-free panels (many
    work-groups, concurrent streams) reproduces the always-GEPP factorisation."""
    say("large speculation vs gepp")
    v, t = meshgen.icosphere(32)
    phi = -1.0 + 0.02 * meshgen.random_sh_field(v, 1, 8, 0)
    c1 = bemhip.Context(0)
    p1, st1 = c1.compute_psi(v, t, phi, "lu64")
    pm, stm = c1.compute_psi(v, t, phi, "lu16x3_ir64")
    p16, st16 = c1.compute_psi(v, t, phi, "lu16_ir64")
    c1.close()
    monkeypatch.setenv("BEMHIP_LU_SPECULATE", "0")
    c2 = bemhip.Context(0)
    p2, st2 = c2.compute_psi(v, t, phi, "lu64")
    c2.close()
    say(st1, st2)
    assert st1["panels_pivoted"] == 0 and st2["panels_pivoted"] == st2["panels"]
    assert rel(p1, p2) < 1e-12
    assert rel(pm, p2) < 1e-8
    assert st16["panels_pivoted"] == 0 and rel(p16, p2) < 1e-8
