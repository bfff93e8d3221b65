"""Failure paths of the step on the HIP path, each against the oracle:
primitives already touching (both touch policies), a degenerate triangle and
parallel edges (the fallback over all edges), every growable broad-phase and
contact buffer starting at 16 entries (regrowth), the outer-iteration cap,
and no triangle-triangle intersection after free-running contact steps."""

import numpy as np
import pytest

import closed_forms as cf
import oracle
import pdipc
import scenes
from parity_util import check_step, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", [1, 0])
def test_touching_pair_toi(policy):
    """A simulated triangle lying in the plane of a static triangle, its
    vertices exactly on it (dyadic coordinates: the distance evaluates to 0),
    moving down: touch_policy 1 gives TOI 0; 0 is the reference rule (NaN,
    the records ignored), leaving TOI 1."""
    s = cf._scene()
    s["params"]["touch_policy"] = policy
    e = 0.0078125
    scenes.add_cloth(s, np.array([[0.0, 0.0, 0.0], [e, 0.0, 0.0], [0.0, 0.0, e]]), np.array([[0, 2, 1]]),
                     strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=1)
    scenes.add_static(s, np.array([[-0.5, 0.0, -0.5], [0.5, 0.0, -0.5], [0.0, 0.0, 0.5]]), np.array([[0, 2, 1]]),
                      patch_dim=1)
    ref = oracle.Oracle(s)
    sim = pdipc.Simulator.from_scene(s)
    x0 = ref.x()
    dx = np.zeros_like(x0)
    dx[: ref.sim, 1] = -1e-3
    print(f"case: touching pair, touch_policy {policy}", flush=True)
    assert len(sim.debug_broad_phase(x0, dx)) == len(ref.debug_broad_phase(x0, dx))
    got, want = sim.debug_toi(x0, x0 + dx), ref.debug_toi(x0, x0 + dx)
    assert got == want
    assert got == (0.0 if policy == 1 else 1.0)


def _degenerate_scene(h=4e-4):
    s = cf._scene()
    tri = np.array([[-0.005, h, 0.0], [0.005, h, 0.0], [0.0, h, 0.01]])
    scenes.add_cloth(s, tri, np.array([[0, 2, 1]]), strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=1)
    seg = np.array([[-0.01, 0.0, 0.0], [0.0, 0.0, 0.0], [0.01, 0.0, 0.0]])  # a zero-area static triangle
    scenes.add_static(s, seg, np.array([[0, 1, 2]]), patch_dim=1)
    return s


def test_degenerate_triangle_and_parallel_edges():
    """Point-triangle records against a zero-area triangle (the minimum over
    its edges) and edge-edge records of parallel edges, h apart: barrier terms
    and TOI against the oracle."""
    sc = _degenerate_scene()
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    x = ref.x()
    print("case: degenerate triangle + parallel edges, barrier terms", flush=True)
    ref.debug_broad_phase(x, np.zeros_like(x))
    sim.debug_broad_phase(x, np.zeros_like(x))
    d0, r0, a0 = ref.debug_collision(x)
    d1, r1, a1 = sim.debug_collision(x)
    assert a0 == a1 and a0 >= 4  # two vertices over the segment, parallel and slanted edges
    assert rel(d1, d0) < 1e-12 and rel(r1, r0) < 1e-12
    print("case: degenerate triangle + parallel edges, TOI", flush=True)
    dx = np.zeros_like(x)
    dx[: ref.sim, 1] = -2e-3
    ref.debug_broad_phase(x, dx)
    sim.debug_broad_phase(x, dx)
    t0, t1 = ref.debug_toi(x, x + dx), sim.debug_toi(x, x + dx)
    assert t0 < 1.0 and abs(t1 - t0) <= 1e-12 * t0


def test_capacity_regrowth_from_tiny_buffers():
    """Three stacked cloths in contact with every growable buffer (traversal
    segments and work items, patch pairs, candidate records, active barrier
    records) starting at 16 entries: the same steps bit for bit as the default
    capacities."""
    s = scenes.make_scene()
    for k in range(3):
        v, f = scenes.cloth_grid(8, side=0.12, height=-0.196 + 3e-3 * k, jitter=5e-4, seed=11 + k,
                                 center=(4e-3 * k, -3e-3 * k))
        scenes.add_cloth(s, v, f, patch_dim=4)
    gv, gf = scenes.ground_plane(y=-0.2, half=0.5)
    scenes.add_static(s, gv, gf, patch_dim=2)
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        print(f"case: regrowth from 16-entry buffers, precision {prec}", flush=True)
        a = pdipc.Simulator.from_scene(s, precision=prec)
        b = pdipc.Simulator.from_scene(s, precision=prec, capacity_hint=16)
        for k in range(10):
            a.step(1)
            b.step(1)
            assert np.array_equal(a.x(), b.x()), k
            ca, cb = a.counters(), b.counters()
            for key in ("outer_iters", "candidates", "patch_pairs", "active_pairs"):
                assert ca[key] == cb[key], (k, key)
        assert ca["candidates"] > 16 and ca["patch_pairs"] > 16 and ca["active_pairs"] > 16


def test_outer_iteration_cap():
    """With an unreachable residual tolerance every step runs the reference's
    30 outer iterations: the same count and the FP64 bar against the oracle."""
    sc = scenes.cloth_over_sphere(12, 2, ground=True, seed=3, patch_dim=12, sphere_patch_dim=12, height=0.16,
                                  params=dict(residual_tol=1e-30))
    ref = oracle.Oracle(sc)
    for _ in range(6):
        ref.step(1)
    sim = pdipc.Simulator.from_scene(sc)
    for k in range(2):
        print(f"case: outer-iteration cap, step {k}", flush=True)
        x, v = ref.x(), ref.v()
        sim.set_x(x)
        sim.set_v(v)
        ref.step(1)
        sim.step(1)
        assert sim.counters()["outer_iters"] == ref.counters()["outer"] == 30
        check_step(sc, x, v, sim.x(), ref.x(), ref.sim, seed=k)


def test_free_running_contact_leaves_no_intersection():
    """Free-running FP32-split steps (not teacher forced) of stacked cloths
    landing on the ground and on each other: no two triangles intersect at any
    step (FP64 edge-triangle crossing count)."""
    s = scenes.make_scene()
    for k in range(3):
        v, f = scenes.cloth_grid(12, side=0.12, height=-0.192 + 3e-3 * k, jitter=5e-4, seed=21 + k,
                                 center=(4e-3 * k, -3e-3 * k))
        scenes.add_cloth(s, v, f, patch_dim=6)
    gv, gf = scenes.ground_plane(y=-0.2, half=0.5)
    scenes.add_static(s, gv, gf, patch_dim=2)
    sim = pdipc.Simulator.from_scene(s, precision=pdipc.FP32_SPLIT)
    assert sim.count_intersections() == 0
    active = 0
    for k in range(30):
        sim.step(1)
        active = max(active, sim.counters()["active_pairs"])
        if k % 10 == 9:
            print(f"case: free-running stacked cloths, step {k + 1}", flush=True)
        assert sim.count_intersections() == 0, k
    assert active > 0
