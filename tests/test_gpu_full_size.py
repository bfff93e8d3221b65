"""GPU parity at the benchmark sizes: one step of the 256x256 cloth over the
subdivision-6 sphere and ground in contact, one step of the 40^3-cell tet
block (320,000 tets) with a 128x128 cloth, and all 100 steps of the 64x64
cloth over the subdivision-4 sphere, against the oracle from identical
states.  Bars as in test_gpu_parity: patch pairs bit-exact as sets, FP64
positions within 1e-8 relative per step (or 20x the oracle's own sensitivity
at an ill-conditioned state), FP32-split within 1e-5 of the FP64 oracle."""

import numpy as np
import pytest

import oracle
import pdipc
import scenes
from parity_util import check_step, rel

pytestmark = pytest.mark.gpu


def _rows(a):
    a = np.asarray(a)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.fixture(scope="module")
def config1_contact():
    """256x256 cloth released 16.6 cm above the ground-level origin at 0.3 m/s
    downwards: the second step touches the sphere (active barrier pairs)."""
    print("case: config-1 sizes, cloth at 0.166 m, v = -0.3 m/s", flush=True)
    sc = scenes.cloth_over_sphere(256, 6, ground=True, seed=0, patch_dim=256, sphere_patch_dim=256, height=0.166)
    ref = oracle.Oracle(sc)
    v = ref.v()
    v[: ref.sim, 1] = -0.3
    ref.set_v(v)
    ref.step(1)
    return sc, ref


@pytest.mark.timeout(600)
def test_config1_step_in_contact(config1_contact):
    sc, ref = config1_contact
    x, v = ref.x(), ref.v()
    sim = pdipc.Simulator.from_scene(sc)
    sim.set_x(x)
    sim.set_v(v)
    print("case: config-1 FP64 step in contact", flush=True)
    ref.step(1)
    sim.step(1)
    c, o = sim.counters(), ref.counters()
    assert o["active_pairs"] > 0
    assert c["outer_iters"] == o["outer"]
    assert c["candidates"] == o["candidates"]
    assert c["patch_pairs"] == o["patch_pairs"]
    err, sens = check_step(sc, x, v, sim.x(), ref.x(), ref.sim)
    print("config-1 FP64 step error", err, "oracle sensitivity", sens)
    print("case: config-1 patch-pair sets", flush=True)
    assert np.array_equal(_rows(sim.patch_pairs()), _rows(ref.patch_pairs()))
    print("case: config-1 zero-intersection check after the contact step", flush=True)
    assert sim.count_intersections() == 0


@pytest.mark.timeout(600)
def test_config2_tet_block_step():
    print("case: config-2 sizes (320,000 tets + 128x128 cloth), FP64 and FP32-split", flush=True)
    sc = scenes.config2()
    ref = oracle.Oracle(sc)
    x, v = ref.x(), ref.v()
    sim = pdipc.Simulator.from_scene(sc)
    sim32 = pdipc.Simulator.from_scene(sc, precision=pdipc.FP32_SPLIT)
    ref.step(1)
    sim.step(1)
    sim32.step(1)
    c, o = sim.counters(), ref.counters()
    assert c["outer_iters"] == o["outer"]
    assert c["candidates"] == o["candidates"]
    err, sens = check_step(sc, x, v, sim.x(), ref.x(), ref.sim)
    print("config-2 FP64 step error", err, "oracle sensitivity", sens)
    e32 = rel(sim32.x()[: ref.sim], ref.x()[: ref.sim])
    print("config-2 FP32-split step error", e32)
    assert e32 < 1e-5


def test_config0_hundred_steps_teacher_forced():
    """The CPU reference path's scene (64x64 cloth over the subdivision-4
    sphere, no ground) for its full 100 steps: every FP64 step of the HIP path
    from the oracle's state at the FP64 bar, with the oracle's outer-iteration
    and candidate counts."""
    sc = scenes.config0()
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    worst, contact = 0.0, 0
    for k in range(100):
        if k % 20 == 0:
            print(f"case: config-0 FP64 step {k}", flush=True)
        x, v = ref.x(), ref.v()
        sim.set_x(x)
        sim.set_v(v)
        ref.step(1)
        sim.step(1)
        err, _ = check_step(sc, x, v, sim.x(), ref.x(), ref.sim, seed=k)
        worst = max(worst, err)
        c, o = sim.counters(), ref.counters()
        assert c["outer_iters"] == o["outer"], (k, c["outer_iters"], o["outer"])
        assert c["candidates"] == o["candidates"], k
        contact += c["active_pairs"] > 0
    print("config-0: worst FP64 step error", worst, "steps with active barrier pairs", contact)
    assert contact > 0  # the cloth reaches the obstacle within the 100 steps
