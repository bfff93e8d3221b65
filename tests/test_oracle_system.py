"""Pins the oracle's assembly, solver and step against independent numpy
constructions, closed forms and the reference's own fixture files."""

import numpy as np
import pytest

import oracle
import scenes


def _tiny_cloth_scene(n=5, seed=0, jitter=2e-3, statics=False):
    s = scenes.make_scene(dict(kappa=1e3))
    v, f = scenes.cloth_grid(n, side=0.5, height=0.2, jitter=jitter, seed=seed)
    scenes.add_cloth(s, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=n)
    if statics:
        sv, sf = scenes.icosphere(1, radius=0.15, noise=0.015, seed=seed)
        scenes.add_static(s, sv, sf, patch_dim=8)
    return s, v, f


def _numpy_cloth_matrix(v, f, strain, bend, density, dt):
    """sysHess of a cloth from first principles: s * (graph Laplacian, counted
    from both endpoints) + L^T diag(bend * area) L (cotangent Laplace-Beltrami
    divided by the vertex area) + M / dt^2.  Angles via arctan2, not acos."""
    n = len(v)
    area = np.zeros(n)
    cot = {}
    for tri in f:
        a, b, c = v[tri]
        ar = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        area[tri] += ar / 3
        for k in range(3):
            i, j, o = tri[k], tri[(k + 1) % 3], tri[(k + 2) % 3]
            e1, e2 = v[i] - v[o], v[j] - v[o]
            ang = np.arctan2(np.linalg.norm(np.cross(e1, e2)), e1 @ e2)
            key = (min(i, j), max(i, j))
            cot[key] = cot.get(key, 0.0) - 0.5 / np.tan(ang)
    L = np.zeros((n, n))
    G = np.zeros((n, n))
    for (i, j), w in cot.items():
        L[i, j] = w / area[i]
        L[j, i] = w / area[j]
        G[i, i] += strain
        G[j, j] += strain
        G[i, j] -= strain
        G[j, i] -= strain
    L[np.diag_indices(n)] = -L.sum(1)
    return G + L.T @ np.diag(bend * area) @ L + np.diag(area * density / dt ** 2)


def test_system_matrix_matches_numpy_assembly():
    s, v, f = _tiny_cloth_scene(n=6)
    o = oracle.Oracle(s)
    A = o.sys_matrix()
    B = _numpy_cloth_matrix(v, f, 1e4, 1e-2, 0.2, 0.01)
    assert np.abs(A - B).max() < 1e-9 * np.abs(B).max()
    assert np.allclose(A, A.T, rtol=0, atol=1e-9 * np.abs(B).max())


def _dense_cheb(A, diag, b, x0, iters, w=2 / 3, rho=0.9992):
    D = np.diag(diag)
    R = np.diag(np.diag(A)) - A  # -offdiag
    Di = np.diag(1 / diag)
    M = (1 - w) * np.eye(len(diag)) + w * Di @ R
    M2 = M @ M
    c = w * w * Di @ R @ Di @ b + (2 - w) * w * Di @ b
    pre, cur = x0, M2 @ x0 + c
    r = rho * rho
    W = 2 / (2 - r)
    for _ in range(1, iters):
        nxt = W * (M2 @ cur + c - pre) + pre
        W = 4 / (4 - r * W)
        pre, cur = cur, nxt
    return cur, M, c, D


def test_chebyshev_solve_matches_dense_recurrence():
    s, v, f = _tiny_cloth_scene(n=6, seed=1)
    o = oracle.Oracle(s)
    A = o.sys_matrix()
    rng = np.random.default_rng(0)
    diag = np.diag(A) * (1 + rng.uniform(0, 0.3, len(A)))
    b = rng.normal(0, 1, (len(A), 3))
    x0 = o.x()
    got = o.debug_solve(diag, b, x0)
    want, M, c, _ = _dense_cheb(A, diag, b, x0[: o.sim], 15)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_one_r2_sweep_is_two_jacobi_sweeps():
    s, v, f = _tiny_cloth_scene(n=5, seed=2)
    s["params"]["cheb_iters"] = 1
    o = oracle.Oracle(s)
    A = o.sys_matrix()
    rng = np.random.default_rng(1)
    diag = np.diag(A).copy()
    b = rng.normal(0, 1, (len(A), 3))
    x0 = o.x()
    got = o.debug_solve(diag, b, x0)
    w = 2 / 3
    R = np.diag(diag) - A

    def jac(x):
        return (1 - w) * x + w * (R @ x + b) / diag[:, None]

    want = jac(jac(x0[: o.sim]))
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


def test_chebyshev_converges_to_linear_solve():
    s, v, f = _tiny_cloth_scene(n=5, seed=3)
    s["params"]["cheb_iters"] = 3000
    o = oracle.Oracle(s)
    A = o.sys_matrix()
    b = np.random.default_rng(2).normal(0, 1, (len(A), 3)) * 1e3
    got = o.debug_solve(np.diag(A).copy(), b, o.x())
    want = np.linalg.solve(A, b)
    assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()


def _rot(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@pytest.mark.parametrize("kind", ["cloth", "tets"])
def test_local_step_rest_fixed_point(kind):
    """At any rigid motion of the rest shape every PD projection returns the
    current element, so rhs(x) = K_elastic x with K_elastic = sysHess - M/dt^2."""
    if kind == "cloth":
        s, v, f = _tiny_cloth_scene(n=5, seed=4)
    else:
        s = scenes.make_scene()
        tv, tt = scenes.tet_block(2, size=0.1)
        scenes.add_tet(s, tv, tt, strain=7e4, density=1000.0, patch_dim=2)
    o = oracle.Oracle(s)
    dt = s["params"]["dt"]
    K = o.sys_matrix() - np.diag(o.mass() / dt ** 2)
    x = o.x()
    R = _rot(7)
    x_rig = x.copy()
    x_rig[: o.sim] = x[: o.sim] @ R.T + np.array([0.1, -0.2, 0.3])
    for xx in (x, x_rig):
        rhs = o.debug_local(xx, np.zeros((o.sim, 3)))
        want = K @ xx[: o.sim]
        assert np.abs(rhs - want).max() < 1e-8 * np.abs(want).max()


def test_free_fall_closed_form():
    s, v, f = _tiny_cloth_scene(n=6, seed=5)
    o = oracle.Oracle(s)
    x0 = o.x()
    o.step(1)
    dt, g = s["params"]["dt"], s["params"]["gravity"]
    x1 = o.x()
    assert np.abs(x1[:, [0, 2]] - x0[:, [0, 2]]).max() < 1e-13
    assert np.abs(x1[:, 1] - (x0[:, 1] + g * dt * dt)).max() < 1e-12
    assert np.abs(o.v()[:, 1] - g * dt * (1 - dt)).max() < 1e-10
    assert o.counters()["outer"] == 2  # second outer pass sees |d|^2 = 0


def test_broad_phase_pairs_match_brute_force():
    s, v, f = _tiny_cloth_scene(n=10, seed=6, statics=True)
    o = oracle.Oracle(s)
    rng = np.random.default_rng(3)
    x = o.x()
    dx = np.zeros_like(x)
    dx[: o.sim] = rng.normal(0, 3e-3, (o.sim, 3))
    dx[: o.sim, 1] -= 0.06  # sweep the cloth into the sphere
    o.debug_broad_phase(x, dx)
    got = {tuple(p) for p in o.patch_pairs()}
    faces = o.boundary_faces()
    infl = 0.5 * s["params"]["dhat"]
    boxes = []
    for pf in o.patch_lists():
        pts = np.concatenate([x[faces[pf].reshape(-1)], (x + dx)[faces[pf].reshape(-1)]])
        boxes.append(np.concatenate([pts.min(0) - infl, pts.max(0) + infl]))
    boxes = np.array(boxes)
    want = set()
    for p in range(len(boxes)):
        lo_ok = (boxes[p, 3:] >= boxes[:, :3]).all(1) & (boxes[:, 3:] >= boxes[p, :3]).all(1)
        for q in np.nonzero(lo_ok)[0]:
            if q > p:
                want.add((p, int(q)))
    assert got == want and len(want) > 0


def test_static_pairs_are_inert():
    """Patch pairs whose patches are both static only ever touch rows that the
    step never reads: skipping them leaves the step bit-identical."""
    s, v, f = _tiny_cloth_scene(n=8, seed=7, statics=True)
    s["params"]["kappa"] = 1e3
    s["cloths"][0]["v"] = s["cloths"][0]["v"] - np.array([0.0, 0.025, 0.0])  # just above the sphere top
    res = []
    for skip in (0, 1):
        s2 = dict(s, params=dict(s["params"], skip_static_pairs=skip))
        o = oracle.Oracle(s2)
        o.step(6)
        res.append((o.x(), o.counters()["candidates"]))
    assert res[0][1] > res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


def test_reference_fixtures(golden):
    for name in ("cloth1", "cloth2"):
        s = scenes.make_scene()
        scenes.add_cloth(s, golden[name + "_v"].astype(np.float64), golden[name + "_f"], strain=1e6, bend=25, pos=1e12,
                         density=2500, patch_dim=50, fixed=golden[name + "_fixed"])
        o = oracle.Oracle(s)
        assert np.array_equal(o.model_dat(0, 0), golden[name + "_boundary"])
        assert np.array_equal(o.model_dat(0, 1), golden[name + "_patch"])


def test_float_mimic_close_to_fp64():
    """The reference's own FP32 arithmetic (Real = float): the local step agrees
    with FP64 to ~4e-7 relative, but every 15-sweep solve adds ~6e-7 m, so one
    step (20 solves) drifts by ~2e-5 m and a few steps by ~1e-4 m -- the error
    the FP32-split mode removes."""
    s = scenes.cloth_over_sphere(8, 1, ground=False, seed=8, patch_dim=8, sphere_patch_dim=8, height=0.17)
    a, b = oracle.Oracle(s), oracle.Oracle(s, use_float=True)
    x = a.x()
    init = np.random.default_rng(0).normal(0, 1, (a.sim, 3))
    ra, rb = a.debug_local(x, init), b.debug_local(x, init)
    assert np.abs(ra - rb).max() < 2e-6 * np.abs(ra).max()
    a.step(1)
    b.step(1)
    e1 = np.abs(a.x() - b.x()).max()
    assert 1e-8 < e1 < 1e-4
    a.step(3)
    b.step(3)
    assert np.abs(a.x() - b.x()).max() < 1e-3
