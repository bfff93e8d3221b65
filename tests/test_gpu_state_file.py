"""Full-state checkpoints (pdipc_save_state / pdipc_load_state): a run restarted
from a saved state on a fresh handle continues bit for bit, and files from
another scene, cut short or with trailing bytes are rejected."""

import numpy as np
import pytest

import pdipc
import scenes

pytestmark = pytest.mark.gpu


def _scene(n=12, seed=0):
    return scenes.cloth_over_sphere(n, 2, ground=True, seed=seed, patch_dim=n, sphere_patch_dim=n, height=0.16)


@pytest.mark.parametrize("precision", [pdipc.FP64, pdipc.FP32_SPLIT])
def test_restart_is_bit_identical(tmp_path, precision):
    sc = _scene()
    path = tmp_path / "state.bin"
    a = pdipc.Simulator.from_scene(sc, precision=precision)
    print(f"case: state checkpoint, precision {precision}", flush=True)
    a.run(8)  # into contact with the obstacle
    a.save_state(path)
    a.run(5)
    b = pdipc.Simulator.from_scene(sc, precision=precision)
    b.load_state(path)
    b.run(5)
    assert np.array_equal(a.x(), b.x())
    assert np.array_equal(a.v(), b.v())
    assert a.counters()["candidates"] == b.counters()["candidates"]


def test_rejects_other_scenes_and_damaged_files(tmp_path):
    sc = _scene()
    path = tmp_path / "state.bin"
    a = pdipc.Simulator.from_scene(sc)
    a.run(2)
    a.save_state(path)
    data = path.read_bytes()
    print("case: state file checks", flush=True)
    other = pdipc.Simulator.from_scene(_scene(n=10))
    with pytest.raises(pdipc.PdipcError, match="different scene"):
        other.load_state(path)
    for name, blob, why in (("short.bin", data[:-8], "truncated"), ("long.bin", data + b"\0", "trailing"),
                            ("bad.bin", b"X" + data[1:], "not a state file")):
        p = tmp_path / name
        p.write_bytes(blob)
        with pytest.raises(pdipc.PdipcError, match=why):
            a.load_state(p)
    x = a.x().copy()
    a.load_state(path)  # the rejected loads left the state alone; this one restores it
    assert np.array_equal(a.x(), x)


@pytest.mark.parametrize("as_double", [0, 1])
def test_frame_dump_has_the_eigen_write_binary_layout(tmp_path, as_double):
    """pdipc_save_frame writes what saveSimulator writes (PdIpcSimulator.h:879-895
    through EigenMatrixIO::write_binary, Commom/EigenMatrixIO.h:22-27): the
    Eigen index type (int64) rows = 3 and cols = points, then the 3 x n points
    matrix column by column (x, y, z of point 0, then point 1, ...), in the
    reference's float (or in double)."""
    sc = scenes.cloth_over_sphere(12, 2, ground=True, seed=2, patch_dim=12, sphere_patch_dim=12)
    sim = pdipc.Simulator.from_scene(sc)
    sim.step(3)
    path = str(tmp_path / "frame_0003.dat")
    print(f"case: frame dump, as_double={as_double}", flush=True)
    sim.save_simulator(path, as_double=bool(as_double))
    raw = open(path, "rb").read()
    x = sim.x()
    n = len(x)
    item = 8 if as_double else 4
    assert len(raw) == 16 + 3 * n * item
    rows, cols = np.frombuffer(raw[:16], "<i8")
    assert (rows, cols) == (3, n)
    data = np.frombuffer(raw[16:], "<f8" if as_double else "<f4")
    cm = data.reshape(cols, rows)  # column-major 3 x n: one column (point) after another
    assert np.array_equal(cm, x.astype(np.float64 if as_double else np.float32))
