"""Shared helpers of the GPU parity tests."""

import os

import numpy as np

import oracle


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_sensitivity(scene, x, v, seed=0, scale=3e-14, samples=4):
    """Largest relative change of one oracle step over a few random input
    perturbations of `scale` relative (the GPU's own rounding floor for one
    step): the conditioning of the reference algorithm itself at this state.
    The amplification is direction dependent, hence several samples."""
    a = oracle.Oracle(scene)
    a.set_x(x)
    a.set_v(v)
    a.step(1)
    worst = 0.0
    for k in range(samples):
        rng = np.random.default_rng(1000 * seed + k)
        b = oracle.Oracle(scene)
        b.set_x(x * (1 + rng.normal(0, scale, x.shape)))
        b.set_v(v)
        b.step(1)
        worst = max(worst, rel(b.x()[: a.sim], a.x()[: a.sim]))
    return worst


# Every step that needed the relaxed bar: (test, seed, err, sens).  The
# session summary (conftest.py) prints them, so a run states how often the
# 1e-8 bar was relaxed and by how much.
FALLBACKS = []
CHECKED = [0]  # steps checked by check_step

# The relaxed bar never exceeds this, however ill-conditioned the state: the
# FP32-split bar, two orders below the ~1e-3 relative motion of a step.
RELAXED_CEILING = 1e-5


def check_step(scene, x, v, got_x, want_x, nsim, tol=1e-8, seed=0):
    """FP64 step parity: `tol`, or -- where the reference algorithm is itself
    ill-conditioned at this state -- 20x the oracle's own sensitivity, capped
    at RELAXED_CEILING.  Each relaxed step is recorded in FALLBACKS."""
    err = rel(got_x[:nsim], want_x[:nsim])
    CHECKED[0] += 1
    if err < tol:
        return err, None
    sens = oracle_sensitivity(scene, x, v, seed)
    test = os.environ.get("PYTEST_CURRENT_TEST", "?").split(" ")[0]
    FALLBACKS.append((test, seed, err, sens))
    assert err < 20 * sens, (err, sens)
    assert err < RELAXED_CEILING, (err, sens)
    return err, sens
