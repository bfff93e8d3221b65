"""Pins the oracle's forward-mode dual numbers (the CPU reference derivative
of the step): its values are bit-identical to the FP64 oracle, and its
tangents agree with central differences of the FP64 oracle through contact,
for per-triangle stiffness and initial-velocity directions, on cloth and on
tets (where they run through the Jacobi SVD)."""

import numpy as np

import oracle
import scenes


def _run(sc, v, steps):
    o = oracle.Oracle(sc)
    o.set_v(v)
    o.step(steps)
    return o.x()[: o.sim]


def test_dual_values_bit_identical():
    sc, v0, _ = scenes.config3(0, n_cloth=8, steps=6)
    a = oracle.Oracle(sc)
    a.set_v(v0)
    a.step(6)
    d = oracle.Oracle(sc, dual=True)
    d.dual_set("v", v0)
    d.step(6)
    x, _ = d.dual_get("x")
    assert np.array_equal(x, a.x()[: a.sim])


def test_dual_tangents_match_central_differences_cloth():
    steps = 14
    sc, v0, _ = scenes.config3(1, n_cloth=10, steps=steps)
    K = oracle.dual_dirs()
    rng = np.random.default_rng(11)
    k0 = sc["cloths"][0]["tri_k"]
    dk = rng.normal(size=(K, len(k0))) * k0
    dv = rng.normal(size=(K,) + v0.shape) * 0.05
    d = oracle.Oracle(sc, dual=True, tri_k_tangents={0: dk})
    d.dual_set("v", v0, dv)
    d.step(steps)
    _, xt = d.dual_get("x")
    assert np.abs(xt).max() > 0
    eps = 1e-6
    errs = []
    for q in range(3):
        sp = dict(sc, cloths=[dict(sc["cloths"][0], tri_k=k0 + eps * dk[q])])
        sm = dict(sc, cloths=[dict(sc["cloths"][0], tri_k=k0 - eps * dk[q])])
        fd = (_run(sp, v0 + eps * dv[q], steps) - _run(sm, v0 - eps * dv[q], steps)) / (2 * eps)
        errs.append(np.abs(fd - xt[q]).max() / np.abs(xt[q]).max())
    assert max(errs) < 1e-3, errs
    assert np.median(errs) < 2e-4, errs


def test_dual_tangents_match_central_differences_tets():
    steps = 4
    sc = scenes.config2(cells=2, cloth_n=6)
    ref = oracle.Oracle(sc)
    rng = np.random.default_rng(5)
    v0 = rng.normal(0, 0.2, (ref.sim, 3))
    K = oracle.dual_dirs()
    dv = rng.normal(size=(K, ref.sim, 3)) * 0.2
    d = oracle.Oracle(sc, dual=True)
    d.dual_set("v", v0, dv)
    d.step(steps)
    _, xt = d.dual_get("x")
    eps = 1e-6
    for q in range(2):
        fd = (_run(sc, v0 + eps * dv[q], steps) - _run(sc, v0 - eps * dv[q], steps)) / (2 * eps)
        assert np.abs(fd - xt[q]).max() < 1e-4 * np.abs(xt[q]).max()
