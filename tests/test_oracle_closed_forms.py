"""Closed forms that pin the oracle's barrier projection, ACCD time-of-impact
value and off-rest cloth projections (each configuration is described in
closed_forms.py).  Every expected value comes from the geometry of the case.
test_oracle_mutants.py rebuilds the oracle with one plausible misreading of
the reference at a time -- the mass split ratios swapped, barycentric-weighted
targets, a wrong static collision mass, another step factor, the gap test
moved, an inverted rest-length or curvature ratio -- and checks that these
tests catch each one."""

import numpy as np
import pytest

import closed_forms as cf
import oracle


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _oracle_collision(scene):
    ref = oracle.Oracle(scene)
    x = ref.x()
    ref.debug_broad_phase(x, np.zeros_like(x))
    diag, rhs, active = ref.debug_collision(x)
    return ref, diag, rhs, active


# ------------------------------------------------------- barrier projection
@pytest.mark.parametrize("case", ["point_over_static_triangle", "point_over_simulated_triangle", "crossing_edges"])
def test_barrier_projection_closed_form(case):
    scene, diag_want, rhs_want, info = getattr(cf, case)()
    ref, diag, rhs, active = _oracle_collision(scene)
    assert active == info["records"]
    assert _rel(diag, diag_want) < 1e-13
    assert _rel(rhs, rhs_want) < 1e-13


# ---------------------------------------------------------------- ACCD value
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("d0,down,up,drift", [
    (1e-3, 3e-3, 1e-3, (0.0, 0.0, 0.0)),        # toi = 0.9 d0 / D = 0.225
    (1e-3, 1.5e-3, 5e-4, (2e-3, -1e-3, 4e-3)),  # the same with a common drift: 0.45
    (1e-3, 4e-4, 1e-4, (1e-3, 0.0, -2e-3)),     # 0.9 d0 / D > 1: the first step returns 1
])
def test_accd_head_on_closed_form(kind, d0, down, up, drift):
    make = cf.head_on_point_triangle if kind == 0 else cf.head_on_edges
    x, dx = make(d0, down, up, drift)
    want = cf.head_on_toi(d0, down + up)
    got = oracle.ccd(kind, x, dx)
    assert abs(got - want) <= 1e-12 * want, (got, want)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("step,gap,iters,want_terms", [
    (0.5, 0.01, 10, 6),    # stops when 0.5^7 < 0.01: six terms, 1 - 0.5^6
    (0.5, 1e-4, 5, 5),     # never reaches the gap: the iteration cap, 1 - 0.5^5
    (0.3, 0.2, 10, 4),     # 0.7^5 < 0.2: four terms
])
def test_accd_geometric_series(kind, step, gap, iters, want_terms):
    d0, D = 1e-3, 8e-3
    make = cf.head_on_point_triangle if kind == 0 else cf.head_on_edges
    x, dx = make(d0, 0.75 * D, 0.25 * D, (1e-3, 2e-3, -1e-3))
    want = (1 - (1 - step) ** want_terms) * d0 / D
    assert abs(cf.head_on_toi(d0, D, step, gap, iters) - want) <= 1e-15
    got = oracle.ccd(kind, x, dx, iters=iters, step=step, gap=gap)
    assert abs(got - want) <= 1e-12 * want, (got, want)


# ------------------------------------------------- off-rest cloth projection
def _curved_scene():
    v, f = cf.curved_sheet()
    s = cf._scene()
    import scenes
    scenes.add_cloth(s, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=7)
    return s, v, f


@pytest.mark.parametrize("lam,shift", [(1.3, (0.01, -0.02, 0.03)), (0.7, (0.0, 0.0, 0.0)), (2.0, (0.1, 0.0, 0.0))])
def test_local_step_uniform_scaling_returns_rest_projection(lam, shift):
    """At x = lam * x_rest + t every spring projects to its rest length along
    the scaled edge, s * l0 * d_hat = s (x_i^r - x_j^r), and the curvature
    vector L x = lam L x_rest is rescaled to its rest norm, back to L x_rest: the
    local step returns K_elastic x_rest whatever lam and t (a curved rest sheet,
    so interior curvature vectors are nonzero)."""
    s, v, f = _curved_scene()
    K, L = cf.numpy_elastic_matrix(v, f, 1e4, 1e-2)
    assert np.linalg.norm(L @ v, axis=1)[8] > 1.0  # an interior vertex is curved
    ref = oracle.Oracle(s)
    n = ref.sim
    zero = np.zeros((n, 3))
    want = K @ v
    rest = ref.debug_local(v, zero)
    assert _rel(rest, want) < 1e-9
    got = ref.debug_local(lam * v + np.asarray(shift), zero)
    assert _rel(got, want) < 1e-9
    # an inverted ratio (|d| / l0, |a| / H0) would scale both terms by lam^2
    assert _rel(lam * lam * want, want) > 0.1


def test_local_step_off_rest_is_not_the_rest_projection():
    """Sanity of the scaling test: a non-uniform deformation changes the result."""
    s, v, f = _curved_scene()
    K, _ = cf.numpy_elastic_matrix(v, f, 1e4, 1e-2)
    ref = oracle.Oracle(s)
    x = v.copy()
    x[:, 0] *= 1.2
    got = ref.debug_local(x, np.zeros((ref.sim, 3)))
    assert _rel(got, K @ v) > 1e-3
