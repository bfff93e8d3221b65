"""Diagnostics through the C ABI.  In-step kernel timing
(pdipc_kernel_timing / pdipc_get_kernel_times): timed steps are the plain
steps bit for bit, and the launch counts follow the step's structure (outer
iterations x local/global passes x sweeps).  Minimum-TOI record tracking
(track_toi_argmin): the same steps bit for bit, with the record of every
clamped outer iteration reported."""

import numpy as np
import pytest

import pdipc
import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [pdipc.FP64, pdipc.FP32_SPLIT])
def test_timed_steps_match_plain_steps_and_count_launches(prec):
    sc = scenes.cloth_over_sphere(24, 2, ground=True, seed=4, patch_dim=24, sphere_patch_dim=12, height=0.16)
    print(f"case: kernel timing, precision {prec}", flush=True)
    a = pdipc.Simulator.from_scene(sc, precision=prec, use_graphs=1)
    b = pdipc.Simulator.from_scene(sc, precision=prec, use_graphs=1)
    a.step(5)
    b.set_x(a.x())
    b.set_v(a.v())
    b.kernel_timing(True)
    outer = lg = 0
    for _ in range(4):
        a.step(1)
        b.step(1)
        c = b.counters()
        outer += c["outer_iters"]
        lg += c["lg_iters"]
        assert np.array_equal(a.x(), b.x())
    t = b.kernel_times()
    p = sc["params"]
    K, L = p["cheb_iters"], p["lg_iters"]
    assert lg == outer * L
    sweeps = "sweep_fp32" if prec == pdipc.FP32_SPLIT else "sweep_fp64"
    assert t[sweeps][1] == lg * (K - 1)  # the first iterate comes from the first term / residual pass
    assert t["rhs"][1] == lg and t["project"][1] == lg  # one cloth: curvature + rhs per pass
    assert 0 < t["ccd"][1] <= outer and t["broad_phase"][1] == 4  # one CCD pass per outer iteration with records
    assert t["operators"][1] == 2 * outer
    for name, (us, n) in t.items():
        if n:
            assert 0.0 < us / n < 1e5, (name, us, n)
    b.kernel_timing(False)
    assert b.kernel_times()[sweeps] == (0.0, 0)


def test_toi_argmin_tracking_changes_nothing():
    """Tracking the frozen choices (minimum-TOI record and its ACCD exit path,
    active barrier record set) leaves the step bit for bit as it is, and the
    hashes are reproducible run to run."""
    sc = scenes.cloth_over_sphere(24, 2, ground=True, seed=4, patch_dim=24, sphere_patch_dim=12, height=0.16)
    print("case: TOI argmin tracking", flush=True)
    a = pdipc.Simulator.from_scene(sc)
    b = pdipc.Simulator.from_scene(sc, track_toi_argmin=1)
    c = pdipc.Simulator.from_scene(sc, track_toi_argmin=1)
    clamped = active = 0
    for k in range(12):
        a.step(1)
        b.step(1)
        c.step(1)
        assert np.array_equal(a.x(), b.x()), k
        ca, cb, cc = a.counters(), b.counters(), c.counters()
        assert ca["toi_argmins"] == 0 and ca["active_sets"] == 0
        assert cb["toi_argmins"] == cc["toi_argmins"] and cb["active_sets"] == cc["active_sets"], k
        if cb["toi_min"] < 1.0:
            clamped += 1
            assert cb["toi_argmins"] != 0, k
        if cb["active_pairs"] > 0:
            active += 1
            assert cb["active_sets"] != 0, k
    assert clamped > 0 and active > 0
