"""The zero-intersection check (pdipc_count_intersections) against a brute-force
count of FP64 edge-triangle crossings over all edge/face pairs, on two cloths
that pass through each other and on two that do not."""

import numpy as np
import pytest

import pdipc
import scenes

pytestmark = pytest.mark.gpu


def _orient(a, b, c, d):
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))


def _brute_force(x, faces, edges):
    """Edge (p, q) crosses face (a, b, c): p and q on strictly different sides
    (or one on the plane), and the line through them inside the triangle."""
    E, F = len(edges), len(faces)
    ei = np.repeat(np.arange(E), F)
    fi = np.tile(np.arange(F), E)
    e, f = edges[ei], faces[fi]
    shared = (e[:, :1] == f).any(1) | (e[:, 1:] == f).any(1)
    e, f = e[~shared], f[~shared]
    p, q = x[e[:, 0]], x[e[:, 1]]
    a, b, c = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    s1, s2 = _orient(a, b, c, p), _orient(a, b, c, q)
    straddle = ~(((s1 > 0) & (s2 > 0)) | ((s1 < 0) & (s2 < 0)) | ((s1 == 0) & (s2 == 0)))
    t1, t2, t3 = _orient(p, q, a, b), _orient(p, q, b, c), _orient(p, q, c, a)
    inside = ((t1 >= 0) & (t2 >= 0) & (t3 >= 0)) | ((t1 <= 0) & (t2 <= 0) & (t3 <= 0))
    inside &= ~((t1 == 0) & (t2 == 0) & (t3 == 0))
    return int((straddle & inside).sum())


def _two_cloths(tilt):
    s = scenes.make_scene()
    v1, f1 = scenes.cloth_grid(10, side=0.5, height=0.0, jitter=1e-3, seed=1)
    v2, f2 = scenes.cloth_grid(10, side=0.4, height=0.0, jitter=1e-3, seed=2)
    c, sn = np.cos(tilt), np.sin(tilt)
    y, z = v2[:, 1].copy(), v2[:, 2].copy()
    v2[:, 1], v2[:, 2] = c * y - sn * z + 0.01, sn * y + c * z  # tilt about x
    if tilt == 0.0:
        v2[:, 1] += 0.05  # parallel, 5 cm above
    scenes.add_cloth(s, v1, f1, patch_dim=4)
    scenes.add_cloth(s, v2, f2, patch_dim=4)
    x = np.concatenate([v1, v2])
    faces = np.concatenate([f1, f2 + len(v1)])
    e = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), 1)
    edges = np.unique(e, axis=0)
    return s, x, faces, edges


@pytest.mark.parametrize("tilt", [0.0, 0.6])
def test_crossing_count_matches_brute_force(tilt):
    print(f"case: two 10x10 cloths, tilt {tilt}", flush=True)
    sc, x, faces, edges = _two_cloths(tilt)
    sim = pdipc.Simulator.from_scene(sc)
    want = _brute_force(x, faces, edges)
    got = sim.count_intersections()
    assert got == want
    assert (want > 0) == (tilt != 0.0)
