"""The HIP path on the closed-form configurations of closed_forms.py, through
the C ABI: barrier projection (k_col_*), the ACCD time of impact (k_ccd, with
the reference constants and with other step / gap / iteration settings) and
the local step at uniformly scaled rest shapes (k_cloth_curvature + k_rhs)."""

import numpy as np
import pytest

import closed_forms as cf
import pdipc
import scenes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("case", ["point_over_static_triangle", "point_over_simulated_triangle", "crossing_edges"])
def test_barrier_projection_closed_form(case):
    scene, diag_want, rhs_want, info = getattr(cf, case)()
    print(f"case: barrier projection, {case}", flush=True)
    sim = pdipc.Simulator.from_scene(scene)
    x = sim.x()
    sim.debug_broad_phase(x, np.zeros_like(x))
    diag, rhs, active = sim.debug_collision(x)
    assert active == info["records"]
    assert _rel(diag, diag_want) < 1e-13
    assert _rel(rhs, rhs_want) < 1e-13


@pytest.mark.parametrize("d0,down,step,gap,iters", [
    (1e-3, 4e-3, 0.9, 0.1, 10),    # the reference constants: 0.9 d0 / D
    (1e-3, 5e-4, 0.9, 0.1, 10),    # 0.9 d0 / D > 1: the first step returns 1
    (1e-3, 8e-3, 0.5, 0.01, 10),   # six terms of the geometric series
    (1e-3, 8e-3, 0.5, 1e-4, 5),    # the iteration cap
])
def test_toi_head_on_closed_form(d0, down, step, gap, iters):
    """A simulated triangle at height d0 over a static triangle's interior moves
    straight down by `down`: each of its vertices is a head-on point-triangle
    record; the step's TOI is the closed form scaled by toi_scale when below 1."""
    scene, _, _, _ = cf.point_over_static_triangle(h=d0)
    scene["params"].update(ccd_step=step, ccd_gap=gap, ccd_iters=iters)
    print(f"case: head-on TOI d0 {d0} down {down} step {step} gap {gap} iters {iters}", flush=True)
    sim = pdipc.Simulator.from_scene(scene)
    x0 = sim.x()
    dx = np.zeros_like(x0)
    dx[: sim.sim, 1] = -down
    sim.debug_broad_phase(x0, dx)
    got = sim.debug_toi(x0, x0 + dx)
    t = cf.head_on_toi(d0, down, step, gap, iters)
    want = t * scene["params"]["toi_scale"] if t < 1.0 else 1.0
    assert abs(got - want) <= 1e-12 * want, (got, want)


@pytest.mark.parametrize("lam,shift", [(1.3, (0.01, -0.02, 0.03)), (0.7, (0.0, 0.0, 0.0))])
def test_local_step_uniform_scaling_returns_rest_projection(lam, shift):
    v, f = cf.curved_sheet()
    s = cf._scene()
    scenes.add_cloth(s, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=7)
    K, _ = cf.numpy_elastic_matrix(v, f, 1e4, 1e-2)
    print(f"case: local step at {lam} x_rest + {shift}", flush=True)
    sim = pdipc.Simulator.from_scene(s)
    zero = np.zeros((sim.sim, 3))
    want = K @ v
    assert _rel(sim.debug_local(v, zero), want) < 1e-9
    assert _rel(sim.debug_local(lam * v + np.asarray(shift), zero), want) < 1e-9
