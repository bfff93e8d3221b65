"""Pins the oracle's primitives against mathematics it does not share code with:
LAPACK (numpy) SVDs, dense brute-force sampling of distances, exact cubic
roots for times of impact, closed forms of the barrier.  Each check is chosen
so that a plausible slip (a transposed operand, a dropped term, a wrong sign or
index) fails it.
"""

import numpy as np
import pytest

import oracle


# ----------------------------------------------------------------- SVD / projection
def _ref_projection(F, lo=0.999, hi=1.001):
    """Projection rule of solveVolumetricStrainConstraintKernel (PdIpcGpuFunc.cu:490-551)
    built on LAPACK: rotations U, V with the sign of det F on the last value."""
    U, s, Vt = np.linalg.svd(F)
    V = Vt.T
    if np.linalg.det(U) < 0:
        U[:, 2] *= -1
        s[2] *= -1
    if np.linalg.det(V) < 0:
        V[:, 2] *= -1
        s[2] *= -1
    s = s.copy()
    if s[0] * s[1] * s[2] < 0:
        s[2] = -s[2]
        s = np.sort(s)[::-1]
    s = np.clip(s, lo, hi)
    return U @ np.diag(s) @ V.T


def test_svd_matches_lapack():
    rng = np.random.default_rng(0)
    for k in range(400):
        F = rng.normal(0, 1, (3, 3))
        if k % 3 == 0:
            F = np.eye(3) + rng.normal(0, 1e-3, (3, 3))  # near-rest, nearly repeated values
        if k % 7 == 0:
            F[:, 0] *= -1  # inverted
        U, s, V = oracle.svd(F)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-12)
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-12 * max(1, np.abs(F).max()))
        sl = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.abs(s), sl, atol=1e-12 * sl[0])
        assert s[0] >= s[1] >= abs(s[2]) - 1e-15
        assert np.sign(s[2]) == np.sign(np.linalg.det(F)) or abs(s[2]) < 1e-12


def test_strain_projection_matches_lapack():
    rng = np.random.default_rng(1)
    for k in range(400):
        F = np.eye(3) + rng.normal(0, 0.05 if k % 2 else 0.0008, (3, 3))
        if k % 5 == 0:
            F = rng.normal(0, 1, (3, 3))
        P = oracle.strain_projection(F)
        assert np.allclose(P, _ref_projection(F), atol=1e-11)


def test_strain_projection_special_cases():
    # inside the clamp band the projection is the identity map
    F = np.diag([1.0005, 0.9995, 1.0])
    assert np.allclose(oracle.strain_projection(F), F, atol=1e-14)
    # a pure stretch clamps to the band edge; a rotation survives unchanged
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    P = oracle.strain_projection(R @ np.diag([1.5, 0.5, 1.0]))
    assert np.allclose(P, R @ np.diag([1.001, 0.999, 1.0]), atol=1e-12)
    # reflection (inverted element): the result is the nearest rotation-like map
    P = oracle.strain_projection(np.diag([1.0, 1.0, -1.0]))
    assert np.allclose(P, np.diag([1.0, 1.0, 1.0]), atol=1e-12)


# ----------------------------------------------------------------- distances
def _pt_brute(p, a, b, c, n=400):
    u = np.linspace(0, 1, n)
    U, V = np.meshgrid(u, u)
    m = U + V <= 1
    q = a[None] + U[m][:, None] * (b - a)[None] + V[m][:, None] * (c - a)[None]
    return ((q - p[None]) ** 2).sum(1).min()


def _ee_brute(a0, a1, b0, b1, n=500):
    r = np.linspace(0, 1, n)
    P = a0[None] + r[:, None] * (a1 - a0)[None]
    Q = b0[None] + r[:, None] * (b1 - b0)[None]
    return ((P[:, None, :] - Q[None, :, :]) ** 2).sum(2).min()


def test_point_triangle_distance():
    rng = np.random.default_rng(2)
    for k in range(60):
        a, b, c = rng.normal(0, 1, (3, 3))
        p = rng.normal(0, 1.5, 3)
        d2, bary = oracle.pt_dist2(p, a, b, c)
        brute = _pt_brute(p, a, b, c)
        assert d2 <= brute + 1e-12
        assert brute - d2 < 2e-3 * max(1.0, brute)
        # the reported barycentric point realises the distance
        q = bary[0] * a + bary[1] * b + bary[2] * c
        assert abs(((p - q) ** 2).sum() - d2) < 1e-12
        assert bary.min() > -1e-12 and abs(bary.sum() - 1) < 1e-12


def test_point_triangle_closed_forms():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    d2, _ = oracle.pt_dist2(np.array([0.2, 0.3, 0.7]), a, b, c)  # above the interior: plane distance
    assert abs(d2 - 0.49) < 1e-15
    d2, bary = oracle.pt_dist2(np.array([-1.0, -2.0, 0.5]), a, b, c)  # vertex region of a
    assert abs(d2 - (1 + 4 + 0.25)) < 1e-14 and abs(bary[0] - 1) < 1e-14
    d2, _ = oracle.pt_dist2(np.array([1.0, 1.0, 0.0]), a, b, c)  # edge region of bc
    assert abs(d2 - 0.5) < 1e-14


def test_edge_edge_distance():
    rng = np.random.default_rng(3)
    for k in range(60):
        a0, a1, b0, b1 = rng.normal(0, 1, (4, 3))
        d2, rs = oracle.ee_dist2(a0, a1, b0, b1)
        brute = _ee_brute(a0, a1, b0, b1)
        assert d2 <= brute + 1e-12
        assert brute - d2 < 2e-3 * max(1.0, brute)
        p = a0 + rs[0] * (a1 - a0)
        q = b0 + rs[1] * (b1 - b0)
        assert abs(((p - q) ** 2).sum() - d2) < 1e-12


def test_edge_edge_closed_forms():
    # skew perpendicular edges crossing at distance 0.3
    d2, rs = oracle.ee_dist2(np.array([-1.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0, -1.0, 0.3]),
                             np.array([0, 1.0, 0.3]))
    assert abs(d2 - 0.09) < 1e-15 and np.allclose(rs, [0.5, 0.5])
    # exactly parallel edges (degenerate cross product): endpoint distance
    d2, _ = oracle.ee_dist2(np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.5, 0.2, 0]),
                            np.array([1.5, 0.2, 0]))
    assert abs(d2 - 0.04) < 1e-15


# ----------------------------------------------------------------- CCD
def _orient(p, a, b, c):
    return np.linalg.det(np.stack([a - p, b - p, c - p]))


def _cubic_roots(f):
    """Real roots in [0, 1] of the cubic t -> f(t) (fit exactly from 4 samples)."""
    ts = np.array([0.0, 1 / 3, 2 / 3, 1.0])
    coef = np.polyfit(ts, [f(t) for t in ts], 3)
    r = np.roots(coef)
    r = r[np.abs(r.imag) < 1e-9].real
    return np.sort(r[(r >= -1e-12) & (r <= 1 + 1e-12)])


def _pt_exact_toi(x, dx):
    p, a, b, c = x
    dp, da, db, dc = dx

    def pos(t):
        return p + t * dp, a + t * da, b + t * db, c + t * dc

    for t in _cubic_roots(lambda t: _orient(*pos(t))):
        P, A, B, C = pos(t)
        d2, _ = oracle.pt_dist2(P, A, B, C)
        if d2 < 1e-18:
            return t
    return None


def test_ccd_point_triangle_conservative():
    rng = np.random.default_rng(4)
    hits = 0
    for k in range(200):
        tri = rng.normal(0, 1e-2, (3, 3))
        cen = tri.mean(0)
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n /= np.linalg.norm(n)
        w = rng.dirichlet([2, 2, 2])
        target = w @ tri
        t_star = rng.uniform(0.1, 0.9)
        vel = n * rng.uniform(2e-3, 2e-2) + rng.normal(0, 2e-3, 3)
        p0 = target - t_star * vel  # passes through `target` at t_star (triangle static)
        x = np.stack([p0, *tri])
        dx = np.zeros((4, 3))
        dx[0] = vel
        if k % 2:  # move everything by a common drift as well
            dx += rng.normal(0, 3e-3, 3)
            x_t = x + t_star * dx
            # keep the crossing: re-aim the point at the moved triangle
            dx[0] = (w @ x_t[1:] - x[0]) / t_star
        te = _pt_exact_toi(x, dx)
        assert te is not None
        toi = oracle.ccd(0, x, dx)
        assert 0 < toi <= te + 1e-12, (k, toi, te)
        # no crossing in [0, toi]: orientation keeps its sign
        s0 = _orient(*x)
        s1 = _orient(*(x + toi * dx))
        assert np.sign(s0) == np.sign(s1)
        hits += 1
    assert hits == 200


def _ee_exact_toi(x, dx):
    def pos(t):
        return x + t * dx

    for t in _cubic_roots(lambda t: _orient(*pos(t))):
        a0, a1, b0, b1 = pos(t)
        d2, _ = oracle.ee_dist2(a0, a1, b0, b1)
        if d2 < 1e-18:
            return t
    return None


def test_ccd_edge_edge_conservative():
    rng = np.random.default_rng(5)
    for k in range(200):
        a0, a1 = rng.normal(0, 1e-2, (2, 3))
        r, s = rng.uniform(0.1, 0.9, 2)
        pa = a0 + r * (a1 - a0)
        dirb = np.cross(a1 - a0, rng.normal(0, 1, 3))
        dirb *= 1e-2 / np.linalg.norm(dirb)
        b0 = pa - s * dirb
        b1 = b0 + dirb  # segment b crosses segment a at pa
        t_star = rng.uniform(0.1, 0.9)
        vel = rng.normal(0, 1e-2, 3)
        x = np.stack([a0, a1, b0 - t_star * vel, b1 - t_star * vel])
        dx = np.zeros((4, 3))
        dx[2:] = vel
        te = _ee_exact_toi(x, dx)
        assert te is not None
        toi = oracle.ccd(1, x, dx)
        assert 0 < toi <= te + 1e-12, (k, toi, te)


def test_ccd_special_cases():
    x = np.array([[0.1, 0.1, 1e-2], [0, 0, 0], [1, 0, 0], [0, 1, 0]], float) * 0.01
    assert oracle.ccd(0, x, np.zeros((4, 3))) == 1.0  # no relative motion
    common = np.tile([1e-3, -2e-3, 5e-4], (4, 1))
    assert oracle.ccd(0, x, common) == 1.0  # rigid common drift
    xt = x.copy()
    xt[0] = [0.001, 0.001, 0.0]  # touching: the point lies on the triangle
    dx = np.zeros((4, 3))
    dx[0, 2] = -1e-3
    assert oracle.ccd(0, xt, dx, touch_policy=1) == 0.0
    assert np.isnan(oracle.ccd(0, xt, dx, touch_policy=0))  # the reference rule: NaN, ignored


# ----------------------------------------------------------------- barrier
def test_barrier_closed_form():
    dh2, kappa = 1e-6, 1e3
    assert oracle.barrier(dh2, dh2, kappa) == 0.0
    assert oracle.barrier(2 * dh2, dh2, kappa) == 0.0
    # at d^2 = dHat^2 / e the log is -1: b = kappa (1 - 1/e)^2
    v = oracle.barrier(dh2 / np.e, dh2, kappa)
    assert abs(v - kappa * (1 - 1 / np.e) ** 2) < 1e-12 * kappa
    ds = np.linspace(1e-3 * dh2, dh2 * 0.999, 200)
    vals = [oracle.barrier(d, dh2, kappa) for d in ds]
    assert all(v > 0 for v in vals)
    assert all(np.diff(vals) < 0)  # decreasing in d^2
