"""Pins of the oracle's time stepping (CPU only): the RK4 step of
LinLinSim::evolve_system_RK4 (LinLinSim.cpp:194-257), the Euler step
(LinLinSim.cpp:260-284), Simulation::get_dt (Simulation.cpp:77-92),
LinLinSim::nopenetration (LinLinSim.cpp:162-190) and the Bernoulli right-hand
side with its pressure forcing (Simulation.cpp:73-75, LinLinSim.cpp:85-146).

The checks come from the mathematics, not from the formulas restated:
orders of convergence measured on the discrete system, closed forms on
hand-made inputs and properties of the exponential wall law.  The mutation
tests at the end build the oracle with planted mistakes (forward Euler,
midpoint RK2, equal stage weights, get_dt without the square root, the wall
law without its entry time) and require that these checks reject every one.
"""

import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import meshgen

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def _case():
    v, t = meshgen.icosphere(4, 1.0)
    phi = -0.1 + 0.05 * meshgen.random_sh_field(v, 1, 3, seed=11)
    # sigma = 0: the curvature term's 6 x 3 FullPivLU fit (LinLinSim.cpp:85-146)
    # changes pivots as the mesh moves, which makes that right-hand side only
    # piecewise smooth in time and hides the order of any integrator
    ph = meshgen.Physics(p_inf=1.0, epsilon=1.0, sigma=0.0, gamma=1.4,
                         field_amp=0.3, field_freq=0.5, field_phase=0.3)
    return v, t, phi, ph


def _run(oracle, lib, method, dt, T, frozen=False):
    v, t, phi, ph = _case()
    p = oracle.make_params(ph, verts=v, tris=t, frozen_stage_time=frozen)
    sim = oracle.Sim(p, v, t, phi, library=lib)
    for _ in range(int(round(T / dt))):
        getattr(sim, method)(dt, True)
    return sim.x.copy(), sim.phi.copy()


def _order_ratios(oracle, lib, method, dts, T, ref_dt):
    xr, pr = _run(oracle, lib, method, ref_dt, T)
    ex, ep = [], []
    for dt in dts:
        x, p = _run(oracle, lib, method, dt, T)
        ex.append(np.max(np.abs(x - xr)))
        ep.append(np.max(np.abs(p - pr)))
    return ex[0] / ex[1], ep[0] / ep[1], ex, ep


def _check_rk4_order(oracle, lib):
    """classical RK4 (stage times t, t+dt/2, t+dt/2, t+dt): halving dt divides
    the error of the discrete system by 2^4"""
    rx, rp, ex, ep = _order_ratios(oracle, lib, "rk4", (0.04, 0.02), 0.16, 0.0025)
    assert 12.0 < rx < 20.0 and 12.0 < rp < 20.0, (rx, rp, ex, ep)


def test_rk4_is_fourth_order(oracle_lib):
    _check_rk4_order(oracle_lib, oracle_lib.lib())


def test_euler_is_first_order(oracle_lib):
    rx, rp, ex, ep = _order_ratios(oracle_lib, oracle_lib.lib(), "euler", (0.01, 0.005), 0.08, 0.000625)
    assert 1.7 < rx < 2.3 and 1.7 < rp < 2.3, (rx, rp, ex, ep)


def _check_frozen_stage_time(oracle, lib):
    """the reference evaluates the Bernoulli right-hand side at the step's
    start time in every stage: with a forcing that vanishes at t = 0 the first
    step equals a step without forcing, which classical stage times do not"""
    v, t, phi, ph = _case()
    ph0 = meshgen.Physics(**{**ph.__dict__, "field_phase": 0.0})
    ph_off = meshgen.Physics(**{**ph.__dict__, "field_amp": 0.0})
    out = {}
    for name, phys, frozen in (("frozen", ph0, True), ("off", ph_off, True), ("classic", ph0, False)):
        p = oracle.make_params(phys, verts=v, tris=t, frozen_stage_time=frozen)
        sim = oracle.Sim(p, v, t, phi, library=lib)
        sim.rk4(0.05, True)
        out[name] = (sim.x.copy(), sim.phi.copy())
    # (not bit-equal: the oracle's threaded assembly sums in a varying order)
    assert np.max(np.abs(out["frozen"][0] - out["off"][0])) < 1e-14
    assert np.max(np.abs(out["frozen"][1] - out["off"][1])) < 1e-14
    assert np.max(np.abs(out["classic"][1] - out["off"][1])) > 1e-4


def test_frozen_stage_time(oracle_lib):
    _check_frozen_stage_time(oracle_lib, oracle_lib.lib())


def _check_get_dt(oracle, lib):
    """dp / (max_i |dpot_i| + dp_balance max_i |v_i|), then min(dt, min_dt)
    when min_dt > 0, on hand-made values whose maxima sit at different
    vertices and where the largest |dpot| is negative"""
    import ctypes as C
    vel = np.array([[3.0, 0.0, 4.0], [1.0, 1.0, 1.0], [-6.0, 0.0, 0.0], [0.0, 2.0, -2.0]])
    dpot = np.array([0.5, -2.5, 1.0, 0.0])
    p = oracle.make_params(meshgen.Physics(), V_0=1.0, dp_balance=3.0)
    f = lambda q: lib.bo_get_dt(C.byref(q), 4, 0.1, np.ascontiguousarray(vel.reshape(-1)), dpot)
    assert abs(f(p) - 0.1 / (2.5 + 3.0 * 6.0)) < 1e-17
    p.dp_balance = 0.5
    assert abs(f(p) - 0.1 / (2.5 + 0.5 * 6.0)) < 1e-17
    p.min_dt = 1e-3
    assert f(p) == 1e-3
    p.min_dt = 1.0
    assert abs(f(p) - 0.1 / (2.5 + 0.5 * 6.0)) < 1e-17


def test_get_dt_closed_form(oracle_lib):
    _check_get_dt(oracle_lib, oracle_lib.lib())


def test_adaptive_step_uses_first_stage_rates(oracle_lib):
    """evolve_system_RK4 with fixdt = false takes get_dt of the first stage's
    velocities and Bernoulli rates"""
    v, t, phi, ph = _case()
    p = oracle_lib.make_params(ph, verts=v, tris=t, frozen_stage_time=True)
    vel, _ = oracle_lib.position_t(p, v, t, phi)
    dpot = oracle_lib.pot_t_multi(p, v, t, vel, 0.0)
    want = oracle_lib.get_dt(p, 0.01, vel, dpot)
    vmax = np.max(np.linalg.norm(vel, axis=1))
    assert abs(want - 0.01 / (np.max(np.abs(dpot)) + 3.0 * vmax)) < 1e-15
    sim = oracle_lib.Sim(p, v, t, phi)
    got = sim.rk4(0.01, False)
    assert abs(got - want) < 1e-15 and sim.time == got


def _wall(lib, eps, dt, x0, cx):
    import oracle
    x = np.array([[x0, 0.2, -0.3]])
    c = np.array([[cx, 0.7, 0.1]])
    xf = np.zeros((1, 3))
    lib.bo_nopenetration(1, float(eps), float(dt), x.reshape(-1), c.reshape(-1), xf.reshape(-1))
    return xf[0], c[0]


def _check_nopenetration(lib):
    eps = 1e-2
    # inside the band: x(dt) = x0 exp(cx dt / x0) solves x' = (cx / x0) x with x'(0) = cx
    x0, cx = 5e-3, -0.5
    xf, c = _wall(lib, eps, 0.1, x0, cx)
    assert abs(xf[0] - x0 * math.exp(-10.0)) < 1e-18
    assert np.allclose(xf[1:], [0.2 + 0.1 * 0.7, -0.3 + 0.1 * 0.1], rtol=0, atol=1e-16)
    assert abs(c[0] - (xf[0] - x0) / 0.1) < 1e-16 and c[1] == 0.7
    h = 1e-7
    slope = (_wall(lib, eps, h, x0, cx)[0][0] - x0) / h
    assert abs(slope - cx) < 1e-4
    for dt in (1e-3, 0.1, 1.0):
        assert _wall(lib, eps, dt, x0, cx)[0][0] > 0.0          # never reaches the wall
    # outside the band: straight motion until x = eps at t_e, then the same
    # law from eps; position and velocity are continuous at t_e
    x0, cx = 5e-2, -0.5
    te = (eps - x0) / cx
    before = _wall(lib, eps, te * (1 - 1e-9), x0, cx)[0][0]
    after = _wall(lib, eps, te * (1 + 1e-9), x0, cx)[0][0]
    assert abs(before - eps) < 1e-10 and abs(after - eps) < 1e-10
    d = 1e-6
    s_after = (_wall(lib, eps, te + 2 * d, x0, cx)[0][0] - _wall(lib, eps, te + d, x0, cx)[0][0]) / d
    assert abs(s_after - cx) < 1e-3, s_after
    dt = te + 0.02
    assert abs(_wall(lib, eps, dt, x0, cx)[0][0] - eps * math.exp(cx / eps * 0.02)) < 1e-15
    assert _wall(lib, eps, 0.5 * te, x0, cx)[0][0] == x0 + 0.5 * te * cx
    # moving away from the wall: plain update
    assert _wall(lib, eps, 0.1, 5e-3, 0.5)[0][0] == 5e-3 + 0.1 * 0.5


def test_nopenetration_wall_law(oracle_lib):
    _check_nopenetration(oracle_lib.lib())


def test_bernoulli_rate_and_forcing(oracle_lib):
    """Dphi/Dt = 2 sigma kappa + |v|^2/2 + p_inf - eps (V_0/V)^gamma + p_field(t),
    p_field = amp sin(2 pi f t + phase): on a sphere of radius R (kappa = 1/R
    there), uniform radial velocity, a chosen V_0 and a forcing that is, at
    f = 0, the same as raising p_inf"""
    R = 1.3
    v, t = meshgen.icosphere(8, R)
    V = oracle_lib.part_volumes(v, t)[0]
    n = v / np.linalg.norm(v, axis=1)[:, None]
    vel = 0.4 * n
    ph = meshgen.Physics(p_inf=0.7, epsilon=1.1, sigma=0.25, gamma=1.4, V_0=2.0 * V)
    p = oracle_lib.make_params(ph)
    got = oracle_lib.pot_t_multi(p, v, t, vel, 0.0)
    want = 2 * 0.25 / R + 0.5 * 0.16 + 0.7 - 1.1 * 2.0 ** 1.4
    assert np.max(np.abs(got - want)) < 2e-3                # kappa is 1/R to O(h^2)
    p0 = oracle_lib.make_params(meshgen.Physics(**{**ph.__dict__, "sigma": 0.0}))
    got0 = oracle_lib.pot_t_multi(p0, v, t, vel, 0.0)
    assert np.max(np.abs(got0 - (0.5 * 0.16 + 0.7 - 1.1 * 2.0 ** 1.4))) < 1e-13
    # a constant forcing is an additive shift of p_inf
    pc = oracle_lib.make_params(meshgen.Physics(**{**ph.__dict__, "field_amp": 0.3, "field_phase": math.pi / 2}))
    ps = oracle_lib.make_params(meshgen.Physics(**{**ph.__dict__, "p_inf": 1.0}))
    assert np.max(np.abs(oracle_lib.pot_t_multi(pc, v, t, vel, 5.0) - oracle_lib.pot_t_multi(ps, v, t, vel, 5.0))) < 1e-15
    # the adjoint workload's p_inf(t) = 1 + 0.3 sin(2 pi 0.5 t)
    pt = oracle_lib.make_params(meshgen.Physics(**{**ph.__dict__, "field_amp": 0.3, "field_freq": 0.5}))
    for tt in (0.37, 1.5, 2.2):
        d = oracle_lib.pot_t_multi(pt, v, t, vel, tt) - got
        assert np.max(np.abs(d - 0.3 * math.sin(math.pi * tt))) < 1e-14


# ------------------------------------------------------------ mutation tests --

MUTANTS = {1: "forward Euler", 2: "midpoint RK2", 3: "equal stage weights", 4: "get_dt without sqrt",
           5: "wall law without entry time"}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory, oracle_lib):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    out = {}
    d = tmp_path_factory.mktemp("mutants")
    for k in MUTANTS:
        so = str(d / f"liboracle_m{k}.so")
        subprocess.run([cc, "-O2", "-fopenmp", "-fPIC", "-std=c11", "-shared", f"-DBO_MUTANT={k}", "-o", so,
                        os.path.join(ORACLE_DIR, "bem_oracle.c"), os.path.join(ORACLE_DIR, "bem_oracle_lu.c"), "-lm"],
                       check=True)
        out[k] = oracle_lib.bind(so)
    return out


@pytest.mark.parametrize("k", [1, 2, 3])
def test_mutant_stage_weights_fail_the_order_check(oracle_lib, mutant_libs, k):
    with pytest.raises(AssertionError):
        _check_rk4_order(oracle_lib, mutant_libs[k])


def test_mutant_get_dt_fails(oracle_lib, mutant_libs):
    with pytest.raises(AssertionError):
        _check_get_dt(oracle_lib, mutant_libs[4])


def test_mutant_wall_law_fails(mutant_libs):
    with pytest.raises(AssertionError):
        _check_nopenetration(mutant_libs[5])
