"""The SELL matrix layout's variants against the oracle: every supported lane
count per row (1, 2, 4; 8 is the default every other test runs), 32-bit
columns (a cloth whose vertex ids are randomly permuted, so that the offsets
col - row of M^2 leave the 16-bit range the FP32 sweep otherwise stores), and
the reordered rows of a layout whose rows differ in width (a tet block: its
rows are sorted by width within windows, so slot and row differ).

Bars as in the parity tests: FP64 within 1e-8 relative per step (teacher
forced), FP32-split within 1e-5 of the FP64 oracle per step."""

import numpy as np
import pytest

import oracle
import pdipc
import scenes
from parity_util import check_step, rel

pytestmark = pytest.mark.gpu


def _teacher_forced(sc, ref, steps, **extra):
    """Steps of the HIP library from the oracle's states; returns the worst
    FP64-bar error (checked inside) or the worst FP32-split relative error."""
    sim = pdipc.Simulator.from_scene(sc, **extra)
    fp64 = extra.get("precision", pdipc.FP64) == pdipc.FP64
    worst = 0.0
    for k in range(steps):
        x, v = ref.x(), ref.v()
        sim.set_x(x)
        sim.set_v(v)
        ref.step(1)
        sim.step(1)
        if fp64:
            err, _ = check_step(sc, x, v, sim.x(), ref.x(), ref.sim, seed=k)
        else:
            err = rel(sim.x()[: ref.sim], ref.x()[: ref.sim])
        worst = max(worst, err)
    return worst


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_lane_counts_match_oracle(lanes):
    sc = scenes.cloth_over_sphere(12, 2, ground=True, seed=3, patch_dim=12, sphere_patch_dim=12, height=0.16)
    start = oracle.Oracle(sc)
    start.step(6)  # into contact with the obstacle
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        print(f"case: {lanes} lanes per SELL row, precision {prec}", flush=True)
        ref = oracle.Oracle(sc)
        ref.set_x(start.x())
        ref.set_v(start.v())
        worst = _teacher_forced(sc, ref, 4, sell_lanes=lanes, precision=prec)
        print("worst step error", worst, flush=True)
        if prec == pdipc.FP32_SPLIT:
            assert worst < 1e-5


def test_32bit_columns_scrambled_cloth():
    n = 190  # 36,100 vertices: a random numbering puts most M^2 offsets past 32767
    sc = scenes.cloth_over_sphere(n, 2, ground=True, seed=4, patch_dim=16, sphere_patch_dim=16, height=0.16)
    cl = sc["cloths"][0]
    perm = np.random.default_rng(4).permutation(len(cl["v"]))
    v = np.empty_like(cl["v"])
    v[perm] = cl["v"]
    cl["v"] = v
    cl["f"] = perm[cl["f"]].astype(np.int32)
    f = cl["f"]
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert np.abs(edges[:, 0] - edges[:, 1]).max() > 32767  # the 1-ring alone is out of 16-bit range
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        print(f"case: scrambled {n}x{n} cloth, precision {prec}", flush=True)
        ref = oracle.Oracle(sc)
        ref.step(2)
        worst = _teacher_forced(sc, ref, 2, precision=prec)
        print("worst step error", worst, flush=True)
        if prec == pdipc.FP32_SPLIT:
            assert worst < 1e-5


@pytest.mark.parametrize("lanes", [8, 4])
def test_reordered_rows_tet_block(lanes):
    sc = scenes.config2(cells=6, cloth_n=16)  # 343 tet + 256 cloth vertices; tet rows 3 to 12 slots wide at 8 lanes
    start = oracle.Oracle(sc)
    start.step(4)
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        print(f"case: tet block, reordered SELL rows, {lanes} lanes, precision {prec}", flush=True)
        ref = oracle.Oracle(sc)
        ref.set_x(start.x())
        ref.set_v(start.v())
        worst = _teacher_forced(sc, ref, 3, sell_lanes=lanes, precision=prec)
        print("worst step error", worst, flush=True)
        if prec == pdipc.FP32_SPLIT:
            assert worst < 1e-5


@pytest.mark.parametrize("which,lanes,stage", [("cloth", 8, 1), ("cloth", 4, 1), ("cloth", 2, 1), ("cloth96", 8, 1),
                                               ("cloth96", 8, 2), ("tets", 8, 1), ("tets", 4, 1)])
def test_staged_sweeps_are_bit_identical(which, lanes, stage):
    """The sweeps that stage their operand in LDS (the default whenever the
    tiles fit) sum every row's entries in the same order as the L2-gather
    sweeps: whole steps agree bit for bit in both precisions, and the staged
    FP64 step meets the oracle's bar.  The 96 x 96 cloth has the uniform
    width 8 (lane-major slices, vector loads; by default its FP32 sweeps take
    the narrow tiles with one-byte indices, with sell_stage = 2 the 16-bit
    tiles); the others are chunk-major."""
    if which == "cloth":
        sc = scenes.cloth_over_sphere(24, 2, ground=True, seed=2, patch_dim=24, sphere_patch_dim=12, height=0.16)
    elif which == "cloth96":
        sc = scenes.cloth_over_sphere(96, 2, ground=True, seed=2, patch_dim=96, sphere_patch_dim=16, height=0.16)
    else:
        sc = scenes.config2(cells=4, cloth_n=8)
    ref = oracle.Oracle(sc)
    for _ in range(6):  # into contact
        ref.step(1)
    probe = pdipc.Simulator.from_scene(sc, init=False, sell_lanes=lanes)
    probe.host_setup()
    assert probe.stage_layout(lanes) is not None
    assert (probe.sell_layout(lanes)["uniform_w"] == 8) == (which == "cloth96")
    for prec in (pdipc.FP64, pdipc.FP32_SPLIT):
        print(f"case: staged ({stage}) vs L2-gather sweeps, {which}, {lanes} lanes, precision {prec}", flush=True)
        sims = [pdipc.Simulator.from_scene(sc, sell_lanes=lanes, precision=prec, sell_stage=st) for st in (stage, 0)]
        assert sims[0].sweep_staging() == (2 if which == "cloth96" and stage == 1 else 1)
        assert sims[1].sweep_staging() == 0
        r = oracle.Oracle(sc)
        r.set_x(ref.x())
        r.set_v(ref.v())
        for k in range(4):
            x, v = r.x(), r.v()
            for sim in sims:
                sim.set_x(x)
                sim.set_v(v)
                sim.step(1)
            r.step(1)
            assert np.array_equal(sims[0].x(), sims[1].x()), (k, rel(sims[0].x(), sims[1].x()))
            assert np.array_equal(sims[0].v(), sims[1].v())
            if prec == pdipc.FP64:
                check_step(sc, x, v, sims[0].x(), r.x(), r.sim, seed=k)
            else:
                assert rel(sims[0].x()[: r.sim], r.x()[: r.sim]) < 1e-5
