"""Shared helpers of the GPU parity tests."""

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


def check_step(scene, x, v, got_x, want_x, nsim, tol=1e-8, seed=0):
    """FP64 step parity: `tol`, or 20x the oracle's own sensitivity where the
    reference algorithm is ill-conditioned at this state."""
    err = rel(got_x[:nsim], want_x[:nsim])
    if err < tol:
        return err, None
    sens = oracle_sensitivity(scene, x, v, seed)
    assert err < 20 * sens, (err, sens)
    return err, sens
