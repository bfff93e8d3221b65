"""Pins the LBVH restatement (oracle/lbvh.py) against closed forms and a
brute-force top-down build; the GPU LBVH is checked against it in
test_gpu_lbvh.py."""

import numpy as np

import lbvh


def _interleave_brute(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_expand_bits_is_bit_interleave():
    rng = np.random.default_rng(0)
    xs = np.concatenate([[0, 1, 2, 511, 512, 1023], rng.integers(0, 1024, 500)])
    for x in xs:
        e = int(lbvh.expand_bits(np.array([x]))[0])
        assert e == _interleave_brute(int(x), 0, 0) >> 2


def test_morton_codes_clamp_and_interleave():
    boxes = np.array([[0, 0, 0, 0, 0, 0], [1, 1, 1, 1, 1, 1], [0.5, 0.25, 0.75, 0.5, 0.25, 0.75]], float)
    codes = lbvh.morton_codes(boxes)
    assert codes[0] == 0
    assert codes[1] == _interleave_brute(1023, 1023, 1023)  # clamped at 1023
    assert codes[2] == _interleave_brute(512, 256, 768)


def _recursive(keys, f, l, out):
    """Top-down build: split [f, l] at the highest bit in which keys[f] and
    keys[l] differ; records (f, l, gamma) for every internal node."""
    if f == l:
        return
    d = (int(keys[f]) ^ int(keys[l])).bit_length() - 1
    gamma = f
    while gamma + 1 <= l and not ((int(keys[gamma + 1]) >> d) & 1):
        gamma += 1
    out.add((f, l, gamma))
    _recursive(keys, f, gamma, out)
    _recursive(keys, gamma + 1, l, out)


def test_karras_matches_top_down_build():
    rng = np.random.default_rng(1)
    for trial in range(6):
        P = [1, 2, 3, 50, 700, 1500][trial]
        boxes = rng.uniform(0, 1, (P, 3))
        if trial >= 4:
            boxes = np.round(boxes * 8) / 8  # many duplicate codes: exercises the index tie-break
        boxes = np.concatenate([boxes, boxes + rng.uniform(0, 0.01, (P, 3))], 1)
        codes = lbvh.morton_codes(boxes)
        keys, order = lbvh.sorted_keys(codes)
        assert np.all(keys[1:] > keys[:-1])  # unique, ascending
        left, right, parent = lbvh.build(keys)
        if P == 1:
            assert len(left) == 0 and parent[0] == -1
            continue
        lo, hi = lbvh.leaf_ranges(left, right, P)
        assert lo[0] == 0 and hi[0] == P - 1
        got = set()
        for n in range(P - 1):
            gamma = hi[left[n]]
            assert lo[right[n]] == gamma + 1
            got.add((int(lo[n]), int(hi[n]), int(gamma)))
            assert parent[left[n]] == n and parent[right[n]] == n
        want = set()
        _recursive(keys, 0, P - 1, want)
        assert got == want
        # every leaf reached exactly once
        leaves = [c for n in range(P - 1) for c in (left[n], right[n]) if c >= P - 1]
        assert sorted(leaves) == list(range(P - 1, 2 * P - 1))
