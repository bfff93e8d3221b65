"""GPU parity: every phase of the step on the HIP library against the oracle on
identical inputs (teacher forcing), then whole steps.

Bars: integer / index work (candidate records, patch pairs) bit-exact as sets;
FP64 positions and velocities within 1e-8 relative per step; FP32-split within
1e-5 of the FP64 oracle per step.
"""

import numpy as np
import pytest

import oracle
import pdipc
import scenes
from parity_util import check_step

pytestmark = pytest.mark.gpu


def _scene(seed=0, n=12, subdiv=2, ground=True, height=0.16, **params):
    return scenes.cloth_over_sphere(n, subdiv, ground=ground, seed=seed, patch_dim=n, sphere_patch_dim=n,
                                    height=height, params=params or None)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _rows(a):
    a = np.asarray(a)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.fixture(scope="module")
def contact_state():
    """A scene a few steps into contact, with the oracle's state."""
    print("case: cloth 12x12 over icosphere(2) + ground, seed 0", flush=True)
    sc = _scene(seed=0)
    ref = oracle.Oracle(sc)
    ref.step(8)
    return sc, ref


def test_local_step(contact_state):
    sc, ref = contact_state
    sim = pdipc.Simulator.from_scene(sc)
    rng = np.random.default_rng(1)
    x = ref.x().copy()
    x[: ref.sim] += rng.normal(0, 1e-4, (ref.sim, 3))
    init = rng.normal(0, 1e-2, (ref.sim, 3))
    print("case: local projection, perturbed state", flush=True)
    got = sim.debug_local(x, init)
    want = ref.debug_local(x, init)
    assert _rel(got, want) < 1e-10


def test_solve(contact_state):
    sc, ref = contact_state
    sim = pdipc.Simulator.from_scene(sc)
    rng = np.random.default_rng(2)
    dt = sc["params"]["dt"]
    diag = ref.elastic_diag() + ref.mass() / dt ** 2 * (1 + rng.uniform(0, 0.5, ref.sim))
    rhs = rng.normal(0, 1.0, (ref.sim, 3))
    x0 = ref.x()
    print("case: Chebyshev solve, random rhs", flush=True)
    got = sim.debug_solve(diag, rhs, x0)
    want = ref.debug_solve(diag, rhs, x0)
    assert _rel(got, want) < 1e-10


@pytest.mark.parametrize("skip_static", [0, 1])
def test_broad_phase_sets(contact_state, skip_static):
    sc, ref0 = contact_state
    sc = dict(sc, params=dict(sc["params"], skip_static_pairs=skip_static))
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    rng = np.random.default_rng(3)
    xold = ref0.x()
    dx = np.zeros_like(xold)
    dx[: ref.sim] = rng.normal(0, 2e-3, (ref.sim, 3))
    print(f"case: broad phase, skip_static={skip_static}", flush=True)
    got = sim.debug_broad_phase(xold, dx)
    want = ref.debug_broad_phase(xold, dx)
    assert len(got) == len(want)
    assert np.array_equal(_rows(got), _rows(want))
    gp = _rows(sim.patch_pairs())
    op = _rows(ref.patch_pairs())
    assert np.array_equal(gp, op)


def test_toi_and_collision(contact_state):
    sc, ref = contact_state
    sim = pdipc.Simulator.from_scene(sc)
    rng = np.random.default_rng(4)
    xold = ref.x()
    x1 = xold.copy()
    x1[: ref.sim, 1] -= 4e-3  # push everything down into the sphere
    x1[: ref.sim] += rng.normal(0, 5e-4, (ref.sim, 3))
    dx = x1 - xold
    print("case: TOI over a downward sweep", flush=True)
    sim.debug_broad_phase(xold, dx)
    ref.debug_broad_phase(xold, dx)
    tg = sim.debug_toi(xold, x1)
    to = ref.debug_toi(xold, x1)
    assert to < 1.0
    assert abs(tg - to) <= 1e-12 * max(to, 1e-300)
    xc = xold + to * dx
    print("case: barrier constraints at the clamped state", flush=True)
    dg, rg, ng = sim.debug_collision(xc)
    do, ro, no = ref.debug_collision(xc)
    assert ng == no and no > 0
    assert _rel(dg, do) < 1e-10
    assert _rel(rg, ro) < 1e-10


@pytest.mark.parametrize("seed", [0, 1])
def test_step_fp64_teacher_forced(seed):
    sc = _scene(seed=seed)
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    worst = 0.0
    for k in range(14):
        print(f"case: FP64 step {k}, seed {seed}", flush=True)
        x, v = ref.x(), ref.v()
        sim.set_x(x)
        sim.set_v(v)
        ref.step(1)
        sim.step(1)
        ex, sens = check_step(sc, x, v, sim.x(), ref.x(), ref.sim, seed=k)
        worst = max(worst, ex)
        if sens is None:
            ev = float(np.abs(sim.v() - ref.v()).max() / max(np.abs(ref.v()).max(), 1e-3))
            assert ev < 1e-6, (k, ev)
        c, o = sim.counters(), ref.counters()
        assert c["outer_iters"] == o["outer"], (k, c["outer_iters"], o["outer"])
        assert c["candidates"] == o["candidates"]
    print("worst FP64 step error", worst)


def test_step_fp32_split_error():
    sc = _scene(seed=2)
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc, precision=pdipc.FP32_SPLIT)
    worst = 0.0
    for k in range(10):
        print(f"case: FP32-split step {k}", flush=True)
        sim.set_x(ref.x())
        sim.set_v(ref.v())
        ref.step(1)
        sim.step(1)
        worst = max(worst, _rel(sim.x()[: ref.sim], ref.x()[: ref.sim]))
    print("worst FP32-split step error", worst)
    assert worst < 1e-5


def test_free_fall_closed_form():
    """No contact, rest-shape constraints: one step of the PD map from rest with
    v = 0 is the inertial prediction x + dt^2 g exactly up to the solver's
    rounding (the projections return the rest shape, a fixed point)."""
    sc = _scene(seed=5, ground=False, height=1.0)
    sim = pdipc.Simulator.from_scene(sc)
    x0 = sim.x()
    print("case: free fall", flush=True)
    sim.step(1)
    x1 = sim.x()
    dt, g = sc["params"]["dt"], sc["params"]["gravity"]
    n = sim.sim
    assert np.abs(x1[:n, [0, 2]] - x0[:n, [0, 2]]).max() < 1e-12
    assert np.abs(x1[:n, 1] - (x0[:n, 1] + g * dt * dt)).max() < 1e-9


def test_graph_replay_matches_eager():
    """The local/global loop replayed from a HIP graph (captured once per
    collision flag and precision) gives bit-identical steps; every replay is
    compared, with the step's own host synchronisation between replays."""
    sc = _scene(seed=6)
    eager = pdipc.Simulator.from_scene(sc, use_graphs=0)
    graph = pdipc.Simulator.from_scene(sc, use_graphs=1)
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        eager.set_precision(prec)
        graph.set_precision(prec)
        for k in range(10):
            print(f"case: graph replay step {k}, precision {prec}", flush=True)
            eager.step(1)
            graph.step(1)
            assert np.array_equal(eager.x(), graph.x()), k
            assert np.array_equal(eager.v(), graph.v()), k


def test_static_pair_skip_is_bit_identical():
    sc = _scene(seed=7)
    a = pdipc.Simulator.from_scene(sc, skip_static_pairs=0)
    b = pdipc.Simulator.from_scene(sc, skip_static_pairs=1)
    for k in range(12):
        print(f"case: static-pair skip step {k}", flush=True)
        a.step(1)
        b.step(1)
        assert a.counters()["candidates"] > b.counters()["candidates"]
        assert np.array_equal(a.x(), b.x()), k
