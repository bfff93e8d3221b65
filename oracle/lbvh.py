"""TEST INFRASTRUCTURE ONLY: restatement of the vendored Karras LBVH
(Simulator/Cuda/lbvh/morton_code.cuh, bvh.cuh) in numpy / plain Python, used to
check the GPU LBVH bit for bit.

  morton codes   morton_code.cuh:9-44 (expand_bits; double version) applied to
                 the box centroid (aabb.cuh:135-141) normalised to the union of
                 all boxes (bvh.cuh:197-221, 316-321)
  sort           thrust::stable_sort_by_key on the 32-bit code (bvh.cuh:329-335)
  tie-break      64-bit key (code << 32) | object index (bvh.cuh:339-356);
                 identical topology to the 32-bit codes when those are unique
  topology       determine_range / find_split (bvh.cuh:69-163), node layout of
                 construct_internal_nodes (bvh.cuh:166-193): internal nodes
                 0..P-2, leaf i at P-1+i
"""

import numpy as np


def expand_bits(v):
    v = np.asarray(v, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    v = (v * np.uint64(0x00010001)) & m & np.uint64(0xFF0000FF)
    v = (v * np.uint64(0x00000101)) & m & np.uint64(0x0F00F00F)
    v = (v * np.uint64(0x00000011)) & m & np.uint64(0xC30C30C3)
    v = (v * np.uint64(0x00000005)) & m & np.uint64(0x49249249)
    return v.astype(np.uint32)


def morton_codes(boxes):
    """boxes: (P, 6) [min xyz, max xyz] doubles -> uint32 codes."""
    lo = boxes[:, :3].min(0)
    hi = boxes[:, 3:].max(0)
    c = (boxes[:, 3:] + boxes[:, :3]) * 0.5
    with np.errstate(divide="ignore", invalid="ignore"):
        p = (c - lo) / (hi - lo)
    r = np.fmin(np.fmax(p * 1024.0, 0.0), 1023.0)
    e = expand_bits(r.astype(np.uint32))
    return (e[:, 0].astype(np.uint64) * 4 + e[:, 1].astype(np.uint64) * 2 + e[:, 2].astype(np.uint64)).astype(np.uint32)


def sorted_keys(codes):
    order = np.argsort(codes, kind="stable")
    keys = (codes[order].astype(np.uint64) << np.uint64(32)) | order.astype(np.uint64)
    return keys, order


def _clz64(x):
    return 64 - int(x).bit_length()


def build(keys):
    """Karras topology over sorted unique 64-bit keys -> (left, right, parent)."""
    P = len(keys)
    keys = [int(k) for k in keys]
    left = np.zeros(max(P - 1, 0), np.int64)
    right = np.zeros(max(P - 1, 0), np.int64)
    parent = np.full(2 * P - 1, -1, np.int64)

    def delta(a, b):
        if b < 0 or b >= P:
            return -1
        return _clz64(keys[a] ^ keys[b])

    for idx in range(P - 1):
        if idx == 0:
            first, last = 0, P - 1
        else:
            L, R = delta(idx, idx - 1), delta(idx, idx + 1)
            d = 1 if R > L else -1
            dmin = min(L, R)
            lmax = 2
            while delta(idx, idx + d * lmax) > dmin:
                lmax <<= 1
            l, t = 0, lmax >> 1
            while t > 0:
                if delta(idx, idx + (l + t) * d) > dmin:
                    l += t
                t >>= 1
            j = idx + l * d
            first, last = (j, idx) if d < 0 else (idx, j)
        if keys[first] == keys[last]:
            gamma = (first + last) >> 1
        else:
            dn = _clz64(keys[first] ^ keys[last])
            split, stride = first, last - first
            while True:
                stride = (stride + 1) >> 1
                mid = split + stride
                if mid < last and _clz64(keys[first] ^ keys[mid]) > dn:
                    split = mid
                if stride <= 1:
                    break
            gamma = split
        lc = gamma + (P - 1 if first == gamma else 0)
        rc = gamma + 1 + (P - 1 if last == gamma + 1 else 0)
        left[idx], right[idx] = lc, rc
        parent[lc] = idx
        parent[rc] = idx
    return left, right, parent


def leaf_ranges(left, right, P):
    """Leaf interval [lo, hi] covered by every node (for invariant checks)."""
    N = 2 * P - 1
    lo = np.zeros(N, np.int64)
    hi = np.zeros(N, np.int64)

    def rec(n):
        if n >= P - 1:
            lo[n] = hi[n] = n - (P - 1)
            return
        rec(left[n])
        rec(right[n])
        lo[n], hi[n] = lo[left[n]], hi[right[n]]

    if P > 0:
        import sys
        sys.setrecursionlimit(max(10000, 4 * P))
        rec(0)
    return lo, hi
