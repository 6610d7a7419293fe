"""The benchmark workloads at their full sizes on the GPU (through the C-ABI).

* one sphere of nu = 64 (N = 40962): every solver against the FP64 LU, the
  analytic sphere solutions and the 4-pi rule of the assembly;
* two coupled spheres of nu = 32 (N = 20484): 10 RK4 steps of 5e-4, mixed
  precision against FP64 from the same initial state;
* the sweep of 256 independent bubbles (nu in {4, 8, 16, 32}): one batched
  compute_psi, mixed against FP64 per bubble, and a few bubbles against their
  own single-bubble solves.

No oracle at these sizes (its LU would take minutes): the checks are the
size-independent ones.  Tolerances: mixed precision 1e-8 relative to the FP64
result (the metric's bar), FP64 paths 1e-10 / 1e-12 as written per check.
"""

import math
import sys

import numpy as np
import pytest

import bemhip
import meshgen

pytestmark = pytest.mark.gpu

MIXED = ("lu16_ir64", "lu16x3_ir64", "lu32_ir64")


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sphere64():
    return meshgen.config_large(seed=0, nu=64)


def test_nu64_solvers_agree(ctx, sphere64):
    c = sphere64
    say("nu=64 lu64")
    p64, st64 = ctx.compute_psi(c.verts, c.tris, c.phi, "lu64")
    assert st64["panels_pivoted"] == 0
    for solver in MIXED:
        say(f"nu=64 {solver}")
        p, st = ctx.compute_psi(c.verts, c.tris, c.phi, solver)
        say(solver, st)
        assert st["converged"] == 1
        assert rel(p, p64) < 1e-8, solver
    # the reference's default solver, with a tolerance it can meet at this size
    say("nu=64 bicgstab64 tol=1e-12")
    pb, stb = ctx.compute_psi(c.verts, c.tris, c.phi, "bicgstab64", tol=1e-12)
    say("bicgstab64", stb)
    assert stb["converged"] == 1 and stb["residual"] <= 1e-12
    assert rel(pb, p64) < 1e-8


def test_nu64_analytic_sphere(ctx, sphere64):
    """uniform phi0 -> psi = -phi0/R; Y_lm -> -(l+1)/R Y_lm (O(h^2): 4.2e-3 at
    nu = 8 for l = 0, /64 at nu = 64); mixed and FP64 agree on both"""
    v, t = sphere64.verts, sphere64.tris
    say("nu=64 analytic")
    one = np.ones(len(v))
    ps, _ = ctx.compute_psi(v, t, one, "lu64")
    assert abs(ps.mean() + 1.0) < 1e-4 and np.std(ps) < 1.5e-5
    Y = meshgen.real_sph_harm(3, v)[(3, -2)]
    py, _ = ctx.compute_psi(v, t, Y, "lu64")
    assert rel(py, -4.0 * Y) < 1e-3
    pm, _ = ctx.compute_psi(v, t, Y, MIXED[0])
    assert rel(pm, py) < 1e-8


def test_nu64_four_pi_rule(ctx, sphere64):
    """every row of H sums to -4 pi (ColocSim.cpp:85-102): H 1 = -4 pi 1"""
    v, t = sphere64.verts, sphere64.tris
    say("nu=64 H 1")
    _, _, h = ctx.assemble(v, t, want_G=False, want_H=False, phi=np.ones(len(v)))
    assert np.max(np.abs(h + 4 * math.pi)) < 1e-10 * 4 * math.pi


def test_two_bubbles_nu32_mixed_vs_fp64(ctx):
    c = meshgen.config_two_bubbles(seed=0, nu=32)
    out = {}
    for solver in ("lu64", MIXED[0]):
        say(f"two bubbles nu=32 {solver}: {c.steps} RK4 steps")
        s = bemhip.ColocSim(ctx, c.verts, c.tris, 1.0, 1.0, 0.0, 1.0, solver=solver)
        s.set_phi(c.phi)
        for _ in range(c.steps):
            s.evolve_system_RK4(c.dt, True)
        out[solver] = s.state()
        s.close()
    x64, phi64, psi64, t64 = out["lu64"]
    x, phi, psi, t = out[MIXED[0]]
    assert abs(t - t64) < 1e-15 and abs(t64 - c.steps * c.dt) < 1e-15
    assert rel(x, x64) < 1e-8 and rel(phi, phi64) < 1e-8 and rel(psi, psi64) < 1e-8
    # the bubbles moved: the comparison is not of two unchanged states
    assert rel(x, c.verts) > 1e-4


def test_batch256_mixed_vs_fp64(ctx):
    cases = meshgen.config_batch(seed=0, count=256)
    for k, cs in enumerate(cases):
        cs.phi = -1.0 + 0.02 * meshgen.random_sh_field(cs.verts, 1, 4, 100 + k)
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    say(f"batch 256: N = {len(v)} lu64")
    p64, st64 = ctx.compute_psi(v, t, phi, "lu64", independent_parts=True)
    say(f"batch 256 {MIXED[0]}")
    pm, stm = ctx.compute_psi(v, t, phi, MIXED[0], independent_parts=True)
    assert stm["converged"] == 1
    for (o, n) in parts:
        assert rel(pm[o:o + n], p64[o:o + n]) < 1e-8
    # spot checks: the largest, a middle and the smallest bubble alone
    sizes = np.array([n for (_, n) in parts])
    for k in (int(np.argmax(sizes)), 128, int(np.argmin(sizes))):
        o, n = parts[k]
        say(f"batch bubble {k} alone (N = {n})")
        one, _ = ctx.compute_psi(cases[k].verts, cases[k].tris, cases[k].phi, "lu64")
        assert rel(p64[o:o + n], one) < 1e-12
