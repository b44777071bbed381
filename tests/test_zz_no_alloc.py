"""bh_direct_noalloc: the direct sum with its staging in the library's fixed
device array instead of per-call allocations.  Its results equal
bh_direct's bit for bit, call after call (the array keeps the last call's
contents; a smaller second call must not see them), and sizes beyond the
array's limit are refused."""
import re
import os

import numpy as np
import pytest

from nemo_amd import engine as E

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_limit_matches_header():
    h = open(os.path.join(ROOT, "include", "bh.h")).read()
    assert int(re.search(r"#define BH_DIRECT_NOALLOC_MAX (\d+)", h).group(1)) == E.DIRECT_NOALLOC_MAX


def test_direct_noalloc_equals_direct_twice_then_smaller():
    e = E.Engine(0)
    try:
        print("no-alloc: 20000-body Plummer, 5000 targets then 700", flush=True)
        m, x, _ = E.mkplummer(20000, seed=8, with_vel=False)
        e.upload(m, x)
        rng = np.random.default_rng(1)
        idx1 = rng.choice(20000, 5000, replace=False)
        p1, a1 = e.direct_noalloc(0.02, idx1)
        q1, b1 = e.direct(0.02, idx1)
        assert np.array_equal(p1, q1) and np.array_equal(a1, b1)
        assert np.isfinite(p1).all() and (p1 < 0).all()
        # second call, smaller input and another system
        m2, x2, _ = E.mkplummer(3000, seed=9, with_vel=False)
        e.upload(m2, x2)
        idx2 = rng.choice(3000, 700, replace=False)
        p2, a2 = e.direct_noalloc(0.05, idx2)
        q2, b2 = e.direct(0.05, idx2)
        assert np.array_equal(p2, q2) and np.array_equal(a2, b2)
        # the largest size the array holds, then one beyond it
        m3, x3, _ = E.mkplummer(E.DIRECT_NOALLOC_MAX + 5, seed=10, with_vel=False)
        e.upload(m3, x3)
        idx3 = np.arange(E.DIRECT_NOALLOC_MAX)
        p3, a3 = e.direct_noalloc(0.01, idx3)
        q3, b3 = e.direct(0.01, idx3)
        assert np.array_equal(p3, q3) and np.array_equal(a3, b3)
        with pytest.raises(E.BHError) as ei:
            e.direct_noalloc(0.01, np.arange(E.DIRECT_NOALLOC_MAX + 1))
        assert ei.value.code == 1
    finally:
        e.close()


def test_direct_noalloc_two_contexts_share_the_device_array():
    """Two contexts on one device, calls in sequence: each sees its own
    results (the array is shared, the calls do not overlap)."""
    e1, e2 = E.Engine(0), E.Engine(0)
    try:
        m, x, _ = E.mkplummer(4000, seed=11, with_vel=False)
        m2, x2, _ = E.mkplummer(2500, seed=12, with_vel=False)
        e1.upload(m, x)
        e2.upload(m2, x2)
        i1, i2 = np.arange(0, 4000, 3), np.arange(0, 2500, 7)
        r1 = e1.direct_noalloc(0.03, i1)
        r2 = e2.direct_noalloc(0.03, i2)
        assert np.array_equal(r1[1], e1.direct(0.03, i1)[1]) and np.array_equal(r2[1], e2.direct(0.03, i2)[1])
    finally:
        e1.close()
        e2.close()
