"""The oracle's parallel loops (solve, operator update, TOI minimum, barrier
activity test) give the serial loops' results bit for bit: FP64 and
forward-mode dual runs through contact with 1 and 4 threads."""

import numpy as np

import oracle
import scenes


def _contact_scene():
    return scenes.cloth_over_sphere(12, 2, ground=True, seed=0, patch_dim=12, sphere_patch_dim=12, height=0.16)


def _run(threads, dual=False):
    oracle.set_threads(threads)
    try:
        sc, v0, _ = scenes.config3(0, n_cloth=8, steps=6)
        if dual:
            K = oracle.dual_dirs()
            rng = np.random.default_rng(5)
            dv = rng.normal(size=(K,) + v0.shape) * 0.05
            o = oracle.Oracle(sc, dual=True)
            o.dual_set("v", v0, dv)
            o.step(6)
            return o.dual_get("x")
        o = oracle.Oracle(_contact_scene())
        o.step(12)
        return o.x(), o.v(), o.counters()
    finally:
        oracle.set_threads(0)


def test_fp64_oracle_thread_count_invariant():
    x1, v1, c1 = _run(1)
    x4, v4, c4 = _run(4)
    assert c1["active_pairs"] > 0 or c1["toi_min"] < 1.0
    assert np.array_equal(x1, x4) and np.array_equal(v1, v4)
    assert c1 == c4


def test_dual_oracle_thread_count_invariant():
    (x1, t1), (x4, t4) = _run(1, dual=True), _run(4, dual=True)
    assert np.array_equal(x1, x4) and np.array_equal(t1, t4)
