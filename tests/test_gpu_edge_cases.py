"""Edge cases of the step against the oracle: the smallest scene (one
triangle, so one SELL slice of mostly padding rows and two or three
broad-phase patches) landing on the ground, and several cloths (models)
stacked a few millimetres apart that land on the ground and then on each other
(contact between simulated models, not only against static meshes).

Bars as in the parity tests: FP64 within 1e-8 relative per step (teacher
forced, with the oracle's outer-iteration and candidate counts), FP32-split
within 1e-5 of the FP64 oracle per step."""

import numpy as np
import pytest

import oracle
import pdipc
import scenes
from parity_util import check_step, rel

pytestmark = pytest.mark.gpu


def _one_triangle():
    s = scenes.make_scene()
    v = np.array([[-0.02, -0.196, -0.01], [0.03, -0.1955, -0.012], [0.004, -0.1965, 0.025]])
    scenes.add_cloth(s, v, np.array([[0, 1, 2]], np.int32), patch_dim=1)
    gv, gf = scenes.ground_plane(y=-0.2, half=0.5)
    scenes.add_static(s, gv, gf, patch_dim=2)
    return s


def _stacked(layers=3, n=8):
    s = scenes.make_scene()
    for k in range(layers):
        v, f = scenes.cloth_grid(n, side=0.12, height=-0.196 + 3e-3 * k, jitter=5e-4, seed=11 + k,
                                 center=(4e-3 * k, -3e-3 * k))
        scenes.add_cloth(s, v, f, patch_dim=4)
    gv, gf = scenes.ground_plane(y=-0.2, half=0.5)
    scenes.add_static(s, gv, gf, patch_dim=2)
    return s


def _teacher_forced_fp64(sc, steps, label):
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    worst, cand = 0.0, 0
    for k in range(steps):
        print(f"case: {label}, FP64 step {k}", flush=True)
        x, v = ref.x(), ref.v()
        sim.set_x(x)
        sim.set_v(v)
        ref.step(1)
        sim.step(1)
        err, _ = check_step(sc, x, v, sim.x(), ref.x(), ref.sim, seed=k)
        worst = max(worst, err)
        c, o = sim.counters(), ref.counters()
        assert c["outer_iters"] == o["outer"], (k, c["outer_iters"], o["outer"])
        assert c["candidates"] == o["candidates"], (k, c["candidates"], o["candidates"])
        cand = max(cand, c["candidates"])
        assert sim.count_intersections() == 0, k  # no two triangles intersect after the step
    print("worst FP64 step error", worst, "most candidates", cand, flush=True)
    return cand


def _fp32_split_worst(sc, steps, label):
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc, precision=pdipc.FP32_SPLIT)
    worst = 0.0
    for k in range(steps):
        print(f"case: {label}, FP32-split step {k}", flush=True)
        sim.set_x(ref.x())
        sim.set_v(ref.v())
        ref.step(1)
        sim.step(1)
        worst = max(worst, rel(sim.x()[: ref.sim], ref.x()[: ref.sim]))
    print("worst FP32-split step error", worst, flush=True)
    return worst


def test_single_triangle_lands_on_ground():
    sc = _one_triangle()
    assert _teacher_forced_fp64(sc, 10, "one triangle on the ground") > 0  # it reaches the ground
    assert _fp32_split_worst(sc, 10, "one triangle on the ground") < 1e-5


def test_stacked_cloths_contact_each_other():
    sc = _stacked()
    assert _teacher_forced_fp64(sc, 12, "three stacked 8x8 cloths") > 0
    assert _fp32_split_worst(sc, 12, "three stacked 8x8 cloths") < 1e-5
