"""Configurations whose barrier projection, ACCD time of impact and off-rest
local step are known in closed form from their geometry, shared by the oracle
tests (CPU) and the HIP-path tests (GPU).

Barrier projection (setCollisionConstraintsKernel, PdIpcGpuFunc.cu:2756-2892):
for an active record at distance d with unit separation direction n (from the
triangle or edge b towards the point or edge a), weight
w = b(d^2) + min(m_p, m_q)/dt^2, with b the IPC barrier (PdIpcGpuFunc.cu:2745-2754),
every one of the four vertices gets w on its diagonal and w * target on its rhs,
where the point / edge-a vertices' target is x + (dHat - d) m_q/(m_p + m_q) n
and the triangle / edge-b vertices' target is x - (dHat - d) m_p/(m_p + m_q) n,
not weighted by barycentric coordinates.  The scenes below place every active
record at a known distance and direction, with every other record of the scene
centimetres outside dHat.

ACCD (ccdVertexTriangleToi / ccdEdgeEdgeToi, PdIpcGpuFunc.cu:2460-2607) on a
head-on approach: the primitives close at a constant relative speed D along
the separation direction (plus a common drift, which the mean-displacement
subtraction removes), so every step of size s * d_k / D leaves
d_{k+1} = (1 - s) d_k and the time of impact is the partial geometric sum
sum_{j<n} s (1 - s)^j d0 / D = (1 - (1 - s)^n) d0 / D, where n is the first
k >= 1 with (1 - s)^(k+1) < g (the step that crosses the gap g d0 is not
added), capped at the iteration count, and 1 once the sum passes 1.
"""

import numpy as np

import scenes

DT = 0.01
DHAT = 1e-3
KAPPA = 1e4
STATIC_MASS = 1e4


def barrier(d2, dhat=DHAT, kappa=KAPPA):
    """The IPC barrier of the reference, -kappa (d^2 - dHat^2)^2 ln(d^2/dHat^2) / dHat^4."""
    dh2 = dhat * dhat
    return -kappa * (d2 - dh2) ** 2 * np.log(d2 / dh2) / (dh2 * dh2)


def _scene():
    return scenes.make_scene(dict(dt=DT, dhat=DHAT, kappa=KAPPA, static_mass=STATIC_MASS))


def point_over_static_triangle(h=4e-4, m=5e3):
    """A small simulated triangle lying flat at height h over the interior of a
    large static triangle in the plane y = 0.  Each of its three vertices is one
    active point-triangle record at distance h, direction +y.  The cloth's
    density makes its vertex mass m (comparable to the static collision mass,
    so the split ratios are far from 0 and 1).  Returns (scene, expected diag,
    expected rhs) over the simulated points."""
    s = _scene()
    tri = np.array([[0.0, h, 0.0], [0.01, h, 0.0], [0.0, h, 0.01]])
    area = 0.5 * 0.01 * 0.01
    scenes.add_cloth(s, tri, np.array([[0, 2, 1]]), strain=1e4, bend=1e-2, pos=1e12, density=3 * m / area,
                     patch_dim=1)
    big = np.array([[-1.0, 0.0, -1.0], [1.0, 0.0, -1.0], [0.0, 0.0, 1.5]])
    scenes.add_static(s, big, np.array([[0, 2, 1]]), patch_dim=1)
    w = barrier(h * h) + min(m, STATIC_MASS) / DT ** 2
    shift = (DHAT - h) * STATIC_MASS / (m + STATIC_MASS)
    diag = np.full(3, w)
    rhs = w * (tri + np.array([0.0, shift, 0.0]))
    return s, diag, rhs, dict(mass=m, records=3)


def point_over_simulated_triangle(h=3e-4, m=2e-5):
    """Two simulated single-triangle cloths of equal vertex mass m: a small one
    flat at height h over the interior of a larger one at y = 0.  Each small
    vertex is one active point-triangle record at distance h, direction +y, so
    the split is 1/2 : 1/2; each large-triangle vertex takes the weight and its
    own target from all three records."""
    s = _scene()
    up = np.array([[0.0, h, 0.0], [0.005, h, 0.0], [0.0, h, 0.005]])
    lo = np.array([[-0.05, 0.0, -0.05], [0.05, 0.0, -0.05], [0.0, 0.0, 0.08]])
    a_up = 0.5 * 0.005 * 0.005
    a_lo = 0.5 * np.linalg.norm(np.cross(lo[1] - lo[0], lo[2] - lo[0]))
    scenes.add_cloth(s, up, np.array([[0, 2, 1]]), strain=1e4, bend=1e-2, pos=1e12, density=3 * m / a_up,
                     patch_dim=1)
    scenes.add_cloth(s, lo, np.array([[0, 2, 1]]), strain=1e4, bend=1e-2, pos=1e12, density=3 * m / a_lo,
                     patch_dim=1)
    w = barrier(h * h) + m / DT ** 2
    half = 0.5 * (DHAT - h)
    diag = np.concatenate([np.full(3, w), np.full(3, 3 * w)])
    rhs = np.concatenate([w * (up + [0.0, half, 0.0]), 3 * w * (lo - [0.0, half, 0.0])])
    return s, diag, rhs, dict(mass=m, records=3)


def crossing_edges(h=5e-4, ma=1.5e-5, mb=4.5e-5):
    """Two simulated single-triangle cloths standing vertically, one in the
    plane z = 0 (its lowest edge along x at height h), one in the plane x = 0
    (its highest edge along z at height 0), so those two edges cross at right
    angles, h apart, at their midpoints: one active edge-edge record with
    direction +y from the lower edge to the upper one; every other primitive
    pair is about a centimetre apart.  Vertex masses ma (upper) and mb (lower)."""
    s = _scene()
    A = np.array([[-0.01, h, 0.0], [0.01, h, 0.0], [0.0, h + 0.02, 0.0]])
    B = np.array([[0.0, 0.0, -0.01], [0.0, 0.0, 0.01], [0.0, -0.02, 0.0]])
    area = 0.5 * 0.02 * 0.02
    scenes.add_cloth(s, A, np.array([[0, 1, 2]]), strain=1e4, bend=1e-2, pos=1e12, density=3 * ma / area,
                     patch_dim=1)
    scenes.add_cloth(s, B, np.array([[0, 2, 1]]), strain=1e4, bend=1e-2, pos=1e12, density=3 * mb / area,
                     patch_dim=1)
    w = barrier(h * h) + min(ma, mb) / DT ** 2
    up = (DHAT - h) * mb / (ma + mb)
    down = (DHAT - h) * ma / (ma + mb)
    diag = np.array([w, w, 0.0, w, w, 0.0])
    rhs = np.zeros((6, 3))
    rhs[0:2] = w * (A[0:2] + [0.0, up, 0.0])
    rhs[3:5] = w * (B[0:2] - [0.0, down, 0.0])
    return s, diag, rhs, dict(records=1)


def head_on_toi(d0, D, step=0.9, gap=0.1, iters=10):
    """The ACCD time of impact of a head-on approach from distance d0 at
    relative speed D (closed form above)."""
    toc = 0.0
    for n in range(1, iters + 1):
        toc = (1.0 - (1.0 - step) ** n) * d0 / D
        if toc > 1.0:
            return 1.0
        if n < iters and (1.0 - step) ** (n + 1) < gap:
            return toc
    return toc


def head_on_point_triangle(d0, down, up, drift):
    """Point above the interior of a triangle in the plane y = 0: the point
    moves down by `down`, the triangle up by `up`, everything by `drift`."""
    x = np.array([[0.002, d0, 0.003], [-0.01, 0.0, -0.01], [0.01, 0.0, -0.01], [0.0, 0.0, 0.012]])
    dx = np.zeros((4, 3))
    dx[0, 1] = -down
    dx[1:, 1] = up
    dx += np.asarray(drift, float)
    return x, dx


def head_on_edges(d0, down, up, drift):
    """Edge a along x at height d0 over edge b along z at y = 0, crossing at
    right angles; a moves down by `down`, b up by `up`, both by `drift`."""
    x = np.array([[-0.01, d0, 0.001], [0.012, d0, 0.001], [0.002, 0.0, -0.011], [0.002, 0.0, 0.009]])
    dx = np.zeros((4, 3))
    dx[:2, 1] = -down
    dx[2:, 1] = up
    dx += np.asarray(drift, float)
    return x, dx


def curved_sheet(n=7, radius=0.3):
    """An n x n cloth grid bent onto a sphere of the given radius (every
    interior vertex has a nonzero discrete mean-curvature vector L x)."""
    v, f = scenes.cloth_grid(n, side=0.2, height=0.0, jitter=0.0, seed=0)
    r2 = v[:, 0] ** 2 + v[:, 2] ** 2
    v[:, 1] = np.sqrt(radius ** 2 - r2) - radius
    return v, f


def numpy_elastic_matrix(v, f, strain, bend):
    """The cloth's elastic PD matrix from first principles: strain * (graph
    Laplacian, counted from both endpoints) + L^T diag(bend * area) L with L
    the cotangent Laplace-Beltrami operator divided by the vertex area
    (angles via arctan2)."""
    n = len(v)
    area = np.zeros(n)
    cot = {}
    for tri in f:
        a, b, c = v[tri]
        area[tri] += 0.5 * np.linalg.norm(np.cross(b - a, c - a)) / 3
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
    return G + L.T @ np.diag(bend * area) @ L, L
