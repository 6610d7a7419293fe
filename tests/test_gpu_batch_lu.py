"""Batched fp16 LU of a batch's size classes (lu16_ir64 with
BEMHIP_INDEPENDENT_PARTS): every bubble of more than 1024 unknowns is
factored in lock step with the other bubbles of its size, solved and refined
with the class's batched triangular solves and residuals.

Parity: part by part the batch must agree with the single-bubble solve of the
same library, with the solver lanes (BEMHIP_BATCH_LU=0) and with the oracle's
FP64 solution to the mixed solvers' 1e-8; the FP64 residual of every bubble
meets the refinement's stopping test; a bubble whose check fails (forced here
with BEMHIP_BATCH_FORCE_REDO) is solved again on its own with the same result.
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


def _class_cases(seed=0):
    """Three bubbles of N = 1442 (nu = 12), two of N = 2562 (nu = 16) and one
    small one (N = 162), uneven radii, jittered, SH-perturbed potentials."""
    specs = [(12, 0.9, 0.2), (16, 1.4, 0.3), (4, 0.6, 0.1), (12, 1.2, 0.4), (16, 0.7, 0.25), (12, 1.0, 0.15)]
    out = []
    for k, (nu, R, A) in enumerate(specs):
        v, t = meshgen.icosphere(nu, R)
        v = v + 1e-3 * R * np.random.default_rng(seed + k).standard_normal(v.shape)
        phi = -0.4 + 0.05 * meshgen.random_sh_field(v, 1, 4, seed + 20 + k)
        out.append(meshgen.Case(f"c{k}", np.ascontiguousarray(v), t, phi,
                                meshgen.Physics(1.0, 1.0, 0.0, 1.0, field_amp=A, field_freq=0.5), dt=1e-2))
    return out


def _context(**env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return bemhip.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def ctx_nobatch():
    c = _context(BEMHIP_BATCH_LU="0")
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_redo():
    c = _context(BEMHIP_BATCH_FORCE_REDO="1")
    yield c
    c.close()


def test_class_lu_matches_lanes_single_and_oracle(ctx, ctx_nobatch, oracle_lib):
    cases = _class_cases()
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    say(f"class LU: N = {len(v)} in {len(cases)} bubbles")
    psi, st = ctx.compute_psi(v, t, phi, "lu16_ir64", independent_parts=True)
    say(f"  {st}")
    assert st["converged"] == 1 and st["iterations"] >= 1 and st["panels_pivoted"] == 0
    # panels: 3 x 2 (N = 1442) + 2 x 3 (N = 2562) with 1024-column blocks
    assert st["panels"] == 12
    psi_l, st_l = ctx_nobatch.compute_psi(v, t, phi, "lu16_ir64", independent_parts=True)
    assert st_l["converged"] == 1
    for k, (c, (o, n)) in enumerate(zip(cases, parts)):
        say(f"  bubble {k}: N = {n}")
        assert rel(psi[o:o + n], psi_l[o:o + n]) < 1e-8
        one, _ = ctx.compute_psi(c.verts, c.tris, c.phi, "lu16_ir64")
        assert rel(psi[o:o + n], one) < 1e-8
        p = oracle_lib.make_params(c.physics, verts=c.verts, tris=c.tris)
        ref = oracle_lib.compute_psi(p, c.verts, c.tris, c.phi)
        assert rel(psi[o:o + n], ref) < 1e-8


def test_class_lu_residuals_meet_the_stopping_test(ctx, oracle_lib):
    """Every bubble's FP64 residual |H phi - G psi|_inf against the oracle's
    G and H is at the refinement's 1e-10 |H phi|_inf (up to the assembly
    difference, 1e-12 of the matrix entries)."""
    cases = _class_cases(seed=7)
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    psi, st = ctx.compute_psi(v, t, phi, "lu16_ir64", independent_parts=True)
    assert st["converged"] == 1
    for k in (0, 1):   # one bubble of each class
        o, n = parts[k]
        G, H = oracle_lib.assemble(cases[k].verts, cases[k].tris)
        b = H @ cases[k].phi
        r = np.max(np.abs(b - G @ psi[o:o + n])) / np.max(np.abs(b))
        say(f"  bubble {k}: residual {r:.2e}")
        assert r < 2e-10


def test_class_lu_forced_redo(ctx, ctx_redo):
    cases = _class_cases(seed=3)
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    psi, st = ctx.compute_psi(v, t, phi, "lu16_ir64", independent_parts=True)
    psi_r, st_r = ctx_redo.compute_psi(v, t, phi, "lu16_ir64", independent_parts=True)
    say(f"  redo: {st_r}")
    # problem 1 of each class was factored again on its own: 2 + 3 more panels
    assert st_r["converged"] == 1 and st_r["panels"] == st["panels"] + 5
    for (o, n) in parts:
        assert rel(psi_r[o:o + n], psi[o:o + n]) < 1e-8


def test_class_lu_rk4_matches_lanes(ctx, ctx_nobatch):
    """Two RK4 steps of the batch with per-part forcing: the class path and the
    solver lanes give the same state."""
    cases = _class_cases(seed=11)
    v, t, phi, parts, amp, freq = meshgen.join_batch(cases)
    out = []
    for c in (ctx, ctx_nobatch):
        s = bemhip.ColocSim(c, v, t, 1.0, 1.0, 0.0, 1.0, solver="lu16_ir64", independent_parts=True)
        s.set_phi(phi)
        s.set_part_physics(field_amp=amp, field_freq=freq)
        s.set_time(0.3)
        for _ in range(2):
            s.evolve_system_RK4(1e-2, True)
        out.append(s.state())
        s.close()
    (x, ph, psi, tm), (xl, phl, psil, tml) = out
    assert tm == tml
    assert rel(x, xl) < 1e-10 and rel(ph, phl) < 1e-8 and rel(psi, psil) < 1e-8


@pytest.fixture(scope="module")
def ctx_graph():
    c = _context(BEMHIP_BATCH_GRAPH="1")
    yield c
    c.close()


@pytest.mark.skip(reason="graph capture of batch stages is switched off: RK4 stepping of a batch with "
                         "BEMHIP_BATCH_GRAPH=1 ended in a host segmentation fault twice (the second run tested "
                         "the lanes' ordering after the graph launch); cause not found")
def test_class_lu_graph_replays_match_eager(ctx_graph, ctx):
    """A batch stage's device work is captured into a graph on the second
    call with the same buffers and replayed after that (a replayed call
    reports ms_assemble = 0, its span in ms_factor).  Every replay equals the
    eager solve exactly, with other device work and host synchronisations
    between the replays, and a replay follows new inputs (set_phi) like the
    eager path."""
    cases = _class_cases(seed=5)
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    other = _class_cases(seed=6)[2]
    s = bemhip.ColocSim(ctx_graph, v, t, 1.0, 1.0, 0.0, 1.0, solver="lu16_ir64", independent_parts=True)
    e = bemhip.ColocSim(ctx, v, t, 1.0, 1.0, 0.0, 1.0, solver="lu16_ir64", independent_parts=True)
    s.set_phi(phi)
    e.set_phi(phi)
    ref = e.compute_psi().copy()
    first = s.compute_psi().copy()
    assert np.array_equal(first, ref)
    for k in range(4):
        ctx_graph.compute_psi(other.verts, other.tris, other.phi, "lu64")   # eager work in between
        s.accum_stats(reset=True)
        got = s.compute_psi().copy()
        st = s.accum_stats(reset=True)
        say(f"  call {k + 2}: {st}")
        assert st["ms_assemble"] == 0.0 and st["ms_factor"] > 0.0   # ran from the graph
        assert np.array_equal(got, ref), rel(got, ref)
    phi2 = phi * 1.1 + 0.01 * meshgen.random_sh_field(v, 1, 3, 77)
    s.set_phi(phi2)
    e.set_phi(phi2)
    ref2 = e.compute_psi().copy()
    got2 = s.compute_psi().copy()
    assert np.array_equal(got2, ref2) and rel(ref2, ref) > 1e-3
    s.close()
    e.close()


@pytest.fixture
def ctx_fresh():
    c = bemhip.Context(0)
    yield c
    c.close()


def test_tuner_batch_keeps_fastest_setting(ctx_fresh):
    """bemhip_sim_tune on a batch: 12 candidate schedules (block 512 / 1024,
    1-3 sub-classes, look-ahead on / off) timed, the fastest kept; the batch's
    solution with the kept setting agrees with the default's to the mixed
    solvers' 1e-8."""
    cases = _class_cases(seed=13)
    v, t, phi, parts, _, _ = meshgen.join_batch(cases)
    s = bemhip.ColocSim(ctx_fresh, v, t, 1.0, 1.0, 0.0, 1.0, solver="lu16_ir64", independent_parts=True)
    s.set_phi(phi)
    before = s.compute_psi().copy()
    table = s.tune()
    for e in table:
        say(f"  {e}")
    assert len(table) == 12 and sum(e["chosen"] for e in table) == 1
    kept = [e for e in table if e["chosen"]][0]
    assert kept["ms"] == min(e["ms"] for e in table) and kept["ms"] > 0.0
    assert {(e["nb"], e["split"], e["lookahead"]) for e in table} == \
        {(nb, sp, la) for nb in (512, 1024) for sp in (1, 2, 3) for la in (0, 1)}
    after = s.compute_psi().copy()
    for (o, n) in parts:
        assert rel(after[o:o + n], before[o:o + n]) < 1e-8
    s.close()


def test_tuner_single_fp16_lu(ctx_fresh, oracle_lib):
    """bemhip_sim_tune on one bubble (N = 2562, lu16_ir64): block 768 / 1024 x
    16 / 32 / 48 compute units for the panel stream; the kept setting solves
    to the oracle's FP64 psi within 1e-8."""
    v, t = meshgen.icosphere(16, 1.0)
    phi = -1.0 + 0.05 * meshgen.random_sh_field(v, 1, 4, 3)
    s = bemhip.ColocSim(ctx_fresh, v, t, 1.0, 1.0, 0.0, 1.0, solver="lu16_ir64")
    s.set_phi(phi)
    table = s.tune()
    assert len(table) == 6 and sum(e["chosen"] for e in table) == 1
    assert {(e["nb"], e["reserve"]) for e in table} == {(nb, r) for nb in (768, 1024) for r in (16, 32, 48)}
    psi = s.compute_psi().copy()
    p = oracle_lib.make_params(meshgen.Physics(1.0, 1.0, 0.0, 1.0), verts=v, tris=t)
    ref = oracle_lib.compute_psi(p, v, t, phi)
    assert rel(psi, ref) < 1e-8
    s.close()
