"""tr_intersect_batch_noalloc (closest hits staged in a static device array, no allocation) against
tr_intersect_batch on the same rays, through the C ABI: two calls in a row, the second smaller, so
the second call finds the first call's rays and hits still in the shared array; plus the size limit.
Hits must be identical (same kernel, same inputs): bit-exact ids and distances."""
import numpy as np
import pytest

from jarabotransientrendering_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3] = rng.uniform(0.02, 0.98, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3] = 1e-5
    r[:, 7] = np.inf
    return r


def _call(fn, h, rays):
    from jarabotransientrendering_amd import _native as nat
    hits = np.full((rays.shape[0], 4), 0xDEADBEEF, dtype=np.uint32)
    nat.check(fn(h, nat.as_ptr(rays), rays.shape[0], nat.as_ptr(hits)))
    return hits


def test_noalloc_matches_allocating_query_on_two_calls():
    from jarabotransientrendering_amd import renderer, _native as nat
    w, *_ = renderer.setup_scene(scenes.icosphere_box(level=4, width=8, height=8))
    lib = nat.lib()
    for k, n in enumerate((nat.TR_NOALLOC_MAX_RAYS, 20011)):
        print(f"noalloc call {k}: {n} rays", flush=True)
        rays = _rays(n, 100 + k)
        ref = _call(lib.tr_intersect_batch, w.handle, rays)
        got = _call(lib.tr_intersect_batch_noalloc, w.handle, rays)
        np.testing.assert_array_equal(got, ref)
        # rays start inside the box: most hit something (the front face is open, so not all do)
        assert (ref[:, 1] != 0xFFFFFFFF).mean() > 0.5


def test_noalloc_rejects_more_than_the_limit():
    from jarabotransientrendering_amd import renderer, _native as nat
    w, *_ = renderer.setup_scene(scenes.cornell_box(width=4, height=4))
    rays = _rays(nat.TR_NOALLOC_MAX_RAYS + 1, 7)
    hits = np.zeros((rays.shape[0], 4), dtype=np.uint32)
    rc = nat.lib().tr_intersect_batch_noalloc(w.handle, nat.as_ptr(rays), rays.shape[0], nat.as_ptr(hits))
    assert rc == nat.TR_ERR_INVALID
    assert b"limit" in nat.lib().tr_last_error()
    assert not hits.any()
    # an empty call is a no-op
    assert nat.lib().tr_intersect_batch_noalloc(w.handle, nat.as_ptr(rays), 0, nat.as_ptr(hits)) == nat.TR_OK
