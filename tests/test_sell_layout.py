"""The SELL layout of the M^2 rows, built on the host (no device): every M^2
entry exactly once, in its row's CSR order, in the slot that the row order
gives it; padding that points at its own row; slice widths that fit their
rows; rows reordered only within two-slice windows and sorted there by width;
16-bit columns equal to col - slot.  The entry count is pinned against the
oracle's M^2 size by the host-operator test; here the layout around it."""

import numpy as np
import pytest

import pdipc
import scenes


def _scene(which):
    if which == "cloth":  # 16 x 16: boundary rows are a large share, widths vary
        return scenes.cloth_over_sphere(16, 2, ground=True, seed=1, patch_dim=16, sphere_patch_dim=16)
    if which == "cloth96":  # 96 x 96: one common slice width costs under 2% more slots
        return scenes.cloth_over_sphere(96, 2, ground=True, seed=1, patch_dim=96, sphere_patch_dim=16)
    return scenes.config2(cells=4, cloth_n=8)


@pytest.mark.parametrize("which,lanes", [("cloth", 8), ("cloth", 2), ("cloth96", 8), ("tets", 8), ("tets", 4),
                                         ("tets", 1)])
def test_sell_layout_holds_every_entry_once(which, lanes):
    sim = pdipc.Simulator.from_scene(_scene(which), init=False)
    sim.host_setup()
    c = sim.counters()
    n, nnz = c["sim_points"], c["nnz_r2"]
    L = sim.sell_layout(lanes)
    S = L["rows_per_slice"]
    assert S == 64 // lanes
    off, w, col, flag = L["slice_off"], L["slice_w"].astype(np.int64), L["col"].astype(np.int64), L["row_flag"]
    ns = len(w)
    assert ns == -(-n // S)
    assert off[0] == 0 and np.array_equal(np.diff(off), w * 64)
    slots = ns * S

    perm = L["perm"]
    if L["uniform_w"]:
        assert perm is None and (w == L["uniform_w"]).all()
    if which == "tets":
        assert perm is not None  # tet rows alternate two widths: never uniform
    if which == "cloth96":
        assert L["uniform_w"] == 8  # the lane-major layout (the others are chunk-major)
    order = np.arange(slots) if perm is None else perm.astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(slots))
    assert np.array_equal(order[n:], np.arange(n, slots))  # padding slots keep their place
    win = 2 * S
    for w0 in range(0, n, win):
        w1 = min(n, w0 + win)
        assert np.array_equal(np.sort(order[w0:w1]), np.arange(w0, w1))  # moves stay inside a window

    e = np.arange(len(col))
    sl = np.searchsorted(off, e, side="right") - 1
    if L["uniform_w"] == 8:  # lane-major slices: a lane's 8 entries are consecutive
        lane_in, chunk = (e - off[sl]) // w[sl], (e - off[sl]) % w[sl]
    else:  # chunk-major: a chunk's 64 lanes are consecutive
        lane_in, chunk = (e - off[sl]) % 64, (e - off[sl]) // 64
    slot = sl * S + lane_in // lanes
    q = chunk * lanes + lane_in % lanes  # position in the row's CSR order
    row = order[slot]
    rr = np.where(row < n, row, n - 1)
    diag = flag < 0
    assert np.array_equal(flag & 0x7FFFFFFF, rr)
    real = (row < n) & ((col != row) | diag)
    assert real.sum() == nnz
    pad = ~real
    assert np.array_equal(col[pad], rr[pad]) and not diag[pad].any()

    cnt = np.bincount(row[real], minlength=n)
    assert cnt.min() >= 1
    assert np.array_equal(np.bincount(row[real & diag], minlength=n), np.ones(n, np.int64))
    assert np.array_equal(col[diag], row[diag])
    key = row[real] * (n + 1) + col[real]
    assert len(np.unique(key)) == nnz  # no entry twice
    # a row's real entries fill its CSR positions 0 .. cnt-1, columns ascending
    o = np.lexsort((q[real], row[real]))
    rs, qs, cs = row[real][o], q[real][o], col[real][o]
    starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    assert np.array_equal(qs, np.arange(nnz) - np.repeat(starts, cnt))
    same = rs[1:] == rs[:-1]
    assert (cs[1:][same] > cs[:-1][same]).all()

    need = -(-cnt // lanes)  # width each row needs, in 64-entry columns
    slot_need = np.zeros(slots, np.int64)
    slot_need[:n] = need[order[:n]]
    fit = slot_need.reshape(ns, S).max(axis=1)
    if L["uniform_w"]:
        assert (w >= fit).all() and (w.sum() - fit.sum()) * 50 <= fit.sum()
    else:
        assert np.array_equal(w, fit)
    if perm is not None:
        for w0 in range(0, n, win):
            seg = slot_need[w0:min(n, w0 + win)]
            assert (np.diff(seg) <= 0).all()  # widest first inside each window

    if L["col16"] is not None:
        assert np.array_equal(L["col16"].astype(np.int64), col - slot)


@pytest.mark.parametrize("which,lanes", [("cloth", 8), ("cloth", 4), ("cloth96", 8), ("cloth96", 2), ("tets", 8),
                                         ("tets", 4)])
def test_stage_tables_map_every_entry_to_its_column(which, lanes):
    """LDS staging: each tile's union is sorted, holds every column of the
    tile's entries (padding included) and nothing else, fits its limits, and
    every entry's union index leads back to its column."""
    sim = pdipc.Simulator.from_scene(_scene(which), init=False)
    sim.host_setup()
    L = sim.sell_layout(lanes)
    T = sim.stage_layout(lanes)
    assert T is not None
    ts = T["tile_slices"]
    off, col = L["slice_off"], L["col"].astype(np.int64)
    ns = len(L["slice_w"])
    nt = -(-ns // ts)
    toff, base, tcol, loc = T["tile_off"], T["tile_base"].astype(np.int64), T["tile_col"].astype(np.int64), T["loc"]
    assert len(toff) == nt + 1 and toff[0] == 0 and len(loc) == len(col)
    assert np.diff(toff).max() == T["stage_max"] <= 1024
    for t in range(nt):
        e0, e1 = off[t * ts], off[min(ns, (t + 1) * ts)]
        u = base[t] + tcol[toff[t]:toff[t + 1]]
        assert np.array_equal(u, np.unique(col[e0:e1]))  # sorted, exactly the tile's columns
        assert tcol[toff[t]] == 0 and u[-1] - u[0] <= 65535
        assert np.array_equal(u[loc[e0:e1].astype(np.int64)], col[e0:e1])


@pytest.mark.parametrize("which", ["cloth", "cloth96", "tets"])
def test_narrow_stage_tables_map_every_entry_to_its_column(which):
    """Narrow staging of the FP32 sweep (8 lanes per row, lane-major layouts
    only): tiles of 2 slices whose unions hold at most 256 columns, sorted and
    exact, every entry's one-byte index leading back to its column."""
    sim = pdipc.Simulator.from_scene(_scene(which), init=False)
    sim.host_setup()
    L = sim.sell_layout(8)
    T = sim.narrow_stage_layout()
    if L["uniform_w"] != 8:
        assert T is None  # chunk-major layouts keep the 16-bit tiles only
        return
    assert T is not None and T["tile_slices"] == 2
    off, col = L["slice_off"], L["col"].astype(np.int64)
    ns = len(L["slice_w"])
    nt = -(-ns // 2)
    toff, base, tcol, loc = T["tile_off"], T["tile_base"].astype(np.int64), T["tile_col"].astype(np.int64), T["loc"]
    assert len(toff) == nt + 1 and toff[0] == 0 and len(loc) == len(col) and loc.dtype == np.uint8
    assert np.diff(toff).max() == T["stage_max"] <= 256
    for t in range(nt):
        e0, e1 = off[t * 2], off[min(ns, (t + 1) * 2)]
        u = base[t] + tcol[toff[t]:toff[t + 1]]
        assert np.array_equal(u, np.unique(col[e0:e1]))
        assert np.array_equal(u[loc[e0:e1].astype(np.int64)], col[e0:e1])


def test_one_lane_per_row_is_not_staged():
    sim = pdipc.Simulator.from_scene(_scene("cloth96"), init=False)
    sim.host_setup()
    assert sim.stage_layout(1) is None
