"""Adjoint of the volume-equivalent radius trajectory (bemhip_sim_adjoint_radius).

The reference has no gradient code, so the gradient is pinned by central
finite differences of a forward model: on a small sphere against the FP64 CPU
oracle's own RK4 (the reference's semantics), and at the adjoint workload
(nu = 16, 50 steps of 1e-2 in p_inf(t) = 1 + 0.3 sin(pi t)) against the
library's FP64 forward path along 20 seeded directions.  Tolerances: 1e-6
relative on the small case, 1e-5 at the full workload for the FP64 and the
mixed-precision adjoint.  L is a sum of squares of small residuals, so the
O(h^2) error of a plain central difference is large relative to L' (for one
step of the small case it shrinks as h^2: 2.6e-2 at h = 1e-3, 2.6e-4 at 1e-4,
2.7e-8 at 1e-6); the tests use Richardson-extrapolated differences (O(h^4)).
"""

import sys

import numpy as np
import pytest

import bemhip
import meshgen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _radii_oracle(oracle_lib, case, V0, x0, phi0, steps):
    p = oracle_lib.make_params(case.physics, V_0=V0)
    sim = oracle_lib.Sim(p, x0, case.tris, phi0)
    out = []
    for _ in range(steps):
        sim.rk4(case.dt, True)
        V = oracle_lib.part_volumes(sim.x, case.tris).sum()
        out.append((3.0 * V / (4.0 * np.pi)) ** (1.0 / 3.0))
    return np.array(out)


def _radii_gpu(sim, tris, x0, phi0, steps, dt):
    sim.set_mesh(x0, tris)
    sim.set_phi(phi0)
    sim.set_time(0.0)
    out = []
    for _ in range(steps):
        sim.evolve_system_RK4(dt, True)
        out.append((3.0 * sim.get_volume() / (4.0 * np.pi)) ** (1.0 / 3.0))
    return np.array(out)


def _central(f, h):
    """Richardson-extrapolated central difference of a scalar function of the
    step: (4 D(h/2) - D(h)) / 3, whose error is O(h^4)."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def test_adjoint_small_vs_oracle_finite_differences(ctx, oracle_lib):
    say("adjoint small vs oracle FD")
    case = meshgen.config_adjoint(nu=4)
    case.dt, steps = 2e-2, 6
    rng = np.random.default_rng(7)
    x0 = case.verts + 1e-2 * rng.standard_normal(case.verts.shape)
    phi0 = -0.2 + 0.05 * meshgen.random_sh_field(x0, 1, 3, 3)
    V0 = float(oracle_lib.part_volumes(case.verts, case.tris).sum())
    target = _radii_oracle(oracle_lib, case, V0, x0, meshgen.adjoint_target_phi(case) - 0.2, steps)
    ph = case.physics
    sim = bemhip.ColocSim(ctx, x0, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, field_amp=ph.field_amp,
                          field_freq=ph.field_freq, solver="lu64", V_0=V0)
    sim.set_phi(phi0)
    L, R, gx, gp = sim.adjoint_radius(steps, case.dt, target, every=2)
    Rref = _radii_oracle(oracle_lib, case, V0, x0, phi0, steps)
    assert np.max(np.abs(R - Rref)) < 1e-12
    # L = sum (R - target)^2: a radius error dR moves L by about 2 sum |R - target| dR
    assert abs(L - np.sum((Rref - target) ** 2)) <= 2e-12 * np.sum(np.abs(Rref - target)) + 1e-24
    for k in range(3):
        dx = rng.standard_normal(x0.shape)
        dp = rng.standard_normal(phi0.shape)
        fd = _central(lambda e: np.sum((_radii_oracle(oracle_lib, case, V0, x0 + e * dx, phi0 + e * dp, steps)
                                        - target) ** 2), 1e-5)
        ad = float(np.sum(gx * dx) + np.sum(gp * dp))
        say(f"  dir {k}: fd {fd:.10e} adjoint {ad:.10e}")
        assert abs(fd - ad) <= 1e-6 * abs(fd)
    # the state is left as it was
    x, phi, _, t = sim.state()
    assert np.array_equal(x, x0) and np.array_equal(phi, phi0) and t == 0.0


def test_adjoint_workload_central_differences(ctx):
    """nu = 16 sphere, 50 steps of 1e-2, checkpoints every 10 steps: the FP64
    and the mixed-precision adjoint against central differences of the FP64
    forward along 20 seeded directions (x0 and phi0 together)."""
    say("adjoint nu=16 50 steps")
    case = meshgen.config_adjoint(seed=0)
    ph = case.physics
    steps, dt = case.steps, case.dt
    x0, phi0 = case.verts, case.phi
    fwd = bemhip.ColocSim(ctx, x0, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, field_amp=ph.field_amp,
                          field_freq=ph.field_freq, solver="lu64")
    V0 = fwd.get_volume()
    target = _radii_gpu(fwd, case.tris, x0, meshgen.adjoint_target_phi(case, seed=1), steps, dt)
    grads = {}
    for solver in ("lu64", "lu16x3_ir64", "lu16_ir64"):
        sim = bemhip.ColocSim(ctx, x0, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma, field_amp=ph.field_amp,
                              field_freq=ph.field_freq, solver=solver, V_0=V0)
        sim.set_phi(phi0)
        L, R, gx, gp = sim.adjoint_radius(steps, dt, target, every=10)
        assert L > 0.0
        grads[solver] = (gx, gp)
        sim.close()
    rng = np.random.default_rng(2024)
    worst = {k: 0.0 for k in grads}
    for k in range(20):
        dx = rng.standard_normal(x0.shape) * 1e-2
        dp = rng.standard_normal(phi0.shape) * 1e-2
        fd = _central(lambda e: np.sum((_radii_gpu(fwd, case.tris, x0 + e * dx, phi0 + e * dp, steps, dt)
                                        - target) ** 2), 1e-3)
        for solver, (gx, gp) in grads.items():
            ad = float(np.sum(gx * dx) + np.sum(gp * dp))
            worst[solver] = max(worst[solver], abs(fd - ad) / abs(fd))
    say(f"  worst relative FD mismatch over 20 directions: {worst}")
    assert worst["lu64"] <= 1e-5 and worst["lu16x3_ir64"] <= 1e-5 and worst["lu16_ir64"] <= 1e-5


def test_adjoint_torch_function(ctx):
    torch = pytest.importorskip("torch")
    import bemhip_torch
    say("adjoint torch wrapper")
    case = meshgen.config_adjoint(nu=4)
    ph = case.physics
    sim = bemhip.ColocSim(ctx, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma,
                          field_amp=ph.field_amp, field_freq=ph.field_freq, solver="lu64")
    target = np.full(4, 1.0)
    x0 = torch.tensor(case.verts, dtype=torch.float64, requires_grad=True)
    phi0 = torch.tensor(case.phi - 0.1, dtype=torch.float64, requires_grad=True)
    loss = bemhip_torch.radius_loss(sim, case.tris, x0, phi0, target, 1e-2, every=2)
    loss.backward()
    L, R, gx, gp = sim.adjoint_radius(4, 1e-2, target, every=2)
    assert abs(loss.item() - L) <= 1e-12 * L
    assert np.max(np.abs(x0.grad.numpy() - gx)) <= 1e-10 * np.max(np.abs(gx))
    assert np.max(np.abs(phi0.grad.numpy() - gp)) <= 1e-10 * np.max(np.abs(gp))


def test_adjoint_rejects_unsupported(ctx):
    case = meshgen.config_adjoint(nu=4)
    sim = bemhip.ColocSim(ctx, case.verts, case.tris, sigma=0.1, solver="lu64")
    with pytest.raises(bemhip.BemHipError):
        sim.adjoint_radius(2, 1e-2, np.ones(2), every=1)
    sim2 = bemhip.ColocSim(ctx, case.verts, case.tris, solver="bicgstab64")
    with pytest.raises(bemhip.BemHipError):
        sim2.adjoint_radius(2, 1e-2, np.ones(2), every=1)


def test_adjoint_transposed_solves_match_library_trsv(monkeypatch):
    """The adjoint's transposed solves G^T lam = b run on the library's own
    persistent triangular solves (U^T then L^T, rows of U^T read as columns of
    the factor) and FP64 transposed residuals.  A context with
    BEMHIP_TRSV_PERSISTENT=0 takes rocBLAS's trsv/gemv for the same solves: the
    gradients of the two agree at nu = 16 (N = 2562: 21 row blocks of 128, the
    last one 2 rows): to 1e-10 relative for the FP64 LU (the forward solves of
    the two contexts differ by rounding only) and to 1e-7 for the fp16 LU (the
    refinement of either stops once its correction is below 1e-9)."""
    say("adjoint transposed solves: own kernels vs rocBLAS trsv")
    case = meshgen.config_adjoint(seed=0)
    ph = case.physics
    steps, dt = 3, case.dt
    target = np.full(steps, 1.0)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BEMHIP_TRSV_PERSISTENT", mode)
        c = bemhip.Context(0)
        try:
            for solver in ("lu64", "lu16_ir64"):
                sim = bemhip.ColocSim(c, case.verts, case.tris, ph.p_inf, ph.epsilon, ph.sigma, ph.gamma,
                                      field_amp=ph.field_amp, field_freq=ph.field_freq, solver=solver)
                sim.set_phi(case.phi)
                L, R, gx, gp = sim.adjoint_radius(steps, dt, target, every=3)
                out[mode, solver] = (L, gx, gp)
                sim.close()
        finally:
            c.close()
    monkeypatch.delenv("BEMHIP_TRSV_PERSISTENT")
    for solver, tl, tg in (("lu64", 1e-12, 1e-10), ("lu16_ir64", 1e-9, 1e-7)):
        L1, gx1, gp1 = out["1", solver]
        L0, gx0, gp0 = out["0", solver]
        ex = np.max(np.abs(gx1 - gx0)) / np.max(np.abs(gx0))
        ep = np.max(np.abs(gp1 - gp0)) / np.max(np.abs(gp0))
        say(f"  {solver}: loss rel {abs(L1 - L0) / L0:.2e}, grad_x rel {ex:.2e}, grad_phi rel {ep:.2e}")
        assert abs(L1 - L0) <= tl * L0
        assert ex <= tg and ep <= tg
