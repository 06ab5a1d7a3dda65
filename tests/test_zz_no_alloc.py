"""nsp_trace_first_hit_noalloc (no device allocation: rays and hits staged
through the library's fixed per-device array) against nsp_trace_first_hit,
which allocates its staging per call: identical indices, distances and
near-boundary flags, on two calls in a row (the second smaller, so the first
call's staged data lies beyond it), then a call of another dimension, and the
size limit refused before any copy."""
import numpy as np
import pytest

from math_amd import native

pytestmark = pytest.mark.gpu


def rays(n, K, seed):
    rng = np.random.default_rng(seed)
    org = rng.normal(size=(K, n))
    org = org / np.linalg.norm(org, axis=1, keepdims=True) * 3
    d = rng.normal(size=(K, n)) * 0.5 - org
    return org.astype(np.float32), d.astype(np.float32)


def test_noalloc_first_hit_equals_allocating_entry_point(ctx):
    v, f = native.meshgen_cube_sphere(4, 6, True, 0.02, 0.0, 1)
    gm = ctx.mesh(v, f)
    for K, seed in ((native.NOALLOC_MAX_RAYS, 1), (5000, 2)):
        print(f"n=4, {K} rays", flush=True)
        org, d = rays(4, K, seed)
        a = gm.first_hit(org, d)
        b = gm.first_hit_noalloc(org, d)
        assert (a[0] >= 0).sum() > K // 10
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    v6, f6 = native.meshgen_cube_sphere(6, 1, True, 0.02, 0.0, 3)
    g6 = ctx.mesh(v6, f6)
    org, d = rays(6, 20000, 3)
    print("n=6, 20000 rays", flush=True)
    for x, y in zip(g6.first_hit(org, d), g6.first_hit_noalloc(org, d)):
        assert np.array_equal(x, y)
    # more rays than the staging holds: refused with a message, nothing launched
    org, d = rays(4, native.NOALLOC_MAX_RAYS + 1, 4)
    with pytest.raises(native.NspError) as e:
        gm.first_hit_noalloc(org, d)
    assert e.value.code == native.NSP_E_INVALID and "at most" in str(e.value)
    # the context works as before afterwards
    idx, _, _ = gm.first_hit_noalloc(np.full((1, 4), 3.0, np.float32), np.full((1, 4), -1.0, np.float32))
    assert idx[0] >= 0
