"""The R-MAT edge stream against CPython's own generator (CPU).

vx.rmat_edges(scale, edgefactor, seed) restates, in C++, the quadrant walk
of scripts/prepare_publication_datasets.py:88-110 over CPython's Mersenne
Twister (random.Random(seed): init_by_array of the seed's 32-bit words,
random() = genrand_res53).  Here the same walk runs on the standard
library's random.Random itself, so both the generator and the walk are
checked against the interpreter that the reference script runs on."""
import random

import numpy as np
import pytest


def stdlib_rmat(scale, edgefactor, seed):
    rng = random.Random(seed)
    n = 1 << scale
    a, b, c = 0.57, 0.19, 0.19
    src, dst = [], []
    for _ in range(edgefactor * n):
        u = v = 0
        step = n >> 1
        while step:
            r = rng.random()
            if r < a:
                pass
            elif r < a + b:
                v += step
            elif r < a + b + c:
                u += step
            else:
                u += step
                v += step
            step >>= 1
        src.append(u)
        dst.append(v)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


@pytest.mark.parametrize("scale,edgefactor,seed", [(1, 4, 0), (4, 16, 1), (7, 8, 2), (9, 16, 12345),
                                                   (10, 2, 2**40 + 7), (11, 1, 1)])
def test_rmat_equals_stdlib_random(scale, edgefactor, seed):
    import velographx_amd as vx

    s, d = vx.rmat_edges(scale, edgefactor, seed)
    ws, wd = stdlib_rmat(scale, edgefactor, seed)
    assert s.dtype == np.uint32 and d.dtype == np.uint32
    assert np.array_equal(s, ws) and np.array_equal(d, wd)


def test_planted_mistake_swapped_quadrants_differs():
    """Swapping the two off-diagonal quadrants (a plausible slip) changes
    the stream, so the comparison above can fail."""
    import velographx_amd as vx

    s, d = vx.rmat_edges(8, 4, 3)
    ws, wd = stdlib_rmat(8, 4, 3)
    assert not np.array_equal(s, wd) or not np.array_equal(d, ws)
