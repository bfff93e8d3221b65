"""The scale run at full size: three stacked 1024 x 1024 cloths (3,145,728
simulated points, 6.36 M triangles) over the sphere and the ground, FP32-split
with FP64 CCD and barrier, 30 steps from rest (the layers reach each other
from about step 25).  The oracle cannot run this size, so the test checks what
must hold whatever the rounding: no triangle-triangle intersection before or
after (FP64 edge-triangle crossings over every edge/face pair of overlapping
patches, pdipc_count_intersections), finite positions, a non-empty candidate
set every step and barrier pairs once the layers touch."""

import time

import numpy as np
import pytest

import pdipc
import scenes

pytestmark = pytest.mark.gpu


def test_config4_full_size_keeps_zero_intersections():
    sc = scenes.config4()
    print("case: config 4 at full size (3 x 1024^2 cloths), 30 FP32-split steps", flush=True)
    t0 = time.time()
    sim = pdipc.Simulator.from_scene(sc, precision=pdipc.FP32_SPLIT, use_graphs=1, skip_static_pairs=1)
    c = sim.counters()
    print(f"  setup {time.time() - t0:.0f} s, nnz(M^2) {c['nnz_r2']}", flush=True)
    assert c["sim_points"] == 3 * 1024 * 1024
    assert sim.count_intersections() == 0
    active_steps = 0
    for k in range(30):
        sim.step(1)
        c = sim.counters()
        assert c["candidates"] > 0 and c["patch_pairs"] > 0, k
        assert 1 <= c["outer_iters"] <= 30, k
        if c["active_pairs"] > 0:
            active_steps += 1
        if k % 5 == 4:
            print(f"  step {k + 1}: outer {c['outer_iters']}, K {c['candidates']}, Q {c['patch_pairs']}, "
                  f"active {c['active_pairs']} ({time.time() - t0:.0f} s)", flush=True)
    x = sim.x()
    assert np.isfinite(x).all()
    assert active_steps > 0
    n = sim.count_intersections()
    print(f"  intersections after 30 steps: {n}", flush=True)
    assert n == 0
