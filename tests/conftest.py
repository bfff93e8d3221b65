import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "pd-ipc-armadillodemo_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch  # noqa: F401  (device counting only)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "cloth_armadillo_fixtures.npz"))


def pytest_terminal_summary(terminalreporter):
    """States how many FP64 steps took the relaxed (sensitivity-scaled) bar."""
    try:
        import parity_util
    except Exception:
        return
    fb = parity_util.FALLBACKS
    n = parity_util.CHECKED[0]
    if not n:
        return
    tr = terminalreporter
    tr.write_line(f"FP64 step parity: {n} step(s) checked, relaxed bar taken on {len(fb)} "
                  f"(ceiling {parity_util.RELAXED_CEILING:g})")
    for test, seed, err, sens in fb:
        tr.write_line(f"  {test} seed/step {seed}: err {err:.3e}, oracle sensitivity {sens:.3e}")
