"""Oracle overlap sets (closed-interval box overlap, bbox.cuh:25-29) against
an O(n^2) brute force on the same float boxes."""
import numpy as np

import oracle as O


def _boxes(n, rng, scale=0.1):
    lo = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    ext = rng.uniform(0, scale, (n, 3)).astype(np.float32)
    # quantise so that many boxes touch exactly
    lo = np.round(lo * 64) / 64
    ext = np.round(ext * 64) / 64
    return np.hstack([lo, lo + ext]).astype(np.float32)


def _brute(A, B, self):
    out = []
    for i in range(len(A)):
        for j in range(len(B)):
            if self and j <= i:
                continue
            if np.all(np.maximum(A[i, :3], B[j, :3]) <= np.minimum(A[i, 3:], B[j, 3:])):
                out.append((i, j))
    return sorted(out)


def test_overlap_pairs_bipartite():
    rng = np.random.default_rng(0)
    A, B = _boxes(200, rng), _boxes(150, rng)
    got = [tuple(p) for p in O.overlap_pairs(A, B).tolist()]
    assert got == _brute(A, B, False)


def test_overlap_pairs_self():
    rng = np.random.default_rng(1)
    A = _boxes(250, rng)
    got = [tuple(p) for p in O.overlap_pairs(A).tolist()]
    assert got == _brute(A, A, True)
