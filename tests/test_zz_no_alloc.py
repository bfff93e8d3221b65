"""pdipc_toi_noalloc (no device allocation; scratch in a library-wide device
array) against pdipc_debug_toi, through the C ABI, on two calls in a row, the
second over fewer candidate records."""

import numpy as np
import pytest
import torch

import pdipc
import scenes

pytestmark = pytest.mark.gpu


def test_toi_noalloc_matches_debug_toi():
    sc = scenes.cloth_over_sphere(16, 2, ground=True, seed=5, patch_dim=16, sphere_patch_dim=12, height=0.16)
    sim = pdipc.Simulator.from_scene(sc)
    sim.run(12)  # the cloth reaches the sphere: candidate records near contact
    x0 = sim.x().copy()
    n = sim.sim
    rng = np.random.default_rng(3)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    counts = []
    for call, scale in enumerate((4e-3, 1.5e-3)):  # the second broad phase is narrower: fewer records
        print(f"case: toi_noalloc call {call}, displacement scale {scale}", flush=True)
        dx = np.zeros_like(x0)
        dx[:n] = rng.normal(0.0, 0.5 * scale, (n, 3)) + np.array([0.0, -2.0 * scale, 0.0])
        recs = sim.debug_broad_phase(x0, dx)
        counts.append(len(recs))
        x1 = x0 + dx
        want = sim.debug_toi(x0, x1)
        a = torch.tensor(x0.reshape(-1), dtype=torch.float64, device=dev)
        b = torch.tensor(x1.reshape(-1), dtype=torch.float64, device=dev)
        out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
        sim.set_x(x0)
        before = sim.x().copy()
        torch.cuda.synchronize(dev)
        sim.toi_noalloc(a.data_ptr(), b.data_ptr(), out.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize(dev)
        got = float(out.item())
        print(f"records {len(recs)}: noalloc {got!r} debug_toi {want!r}", flush=True)
        assert got == want
        assert want < 1.0  # the displacement reaches contact: a real minimum, not the default
        assert np.array_equal(sim.x(), before)  # the state is not touched
    assert counts[1] < counts[0]
