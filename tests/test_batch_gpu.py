"""Batched multi-scene renders, hipGraph replay and the launch-shape autotuner (through the C ABI).
A batched job draws the same random streams as the same render on its own, so its film must equal
the single render's film up to the order of float atomic additions."""
import numpy as np
import pytest

from jarabotransientrendering_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    from jarabotransientrendering_amd import renderer
    return renderer


def _jobs():
    box = scenes.cornell_box(width=12, height=10, time_bins=128, sqrt_spp=3, n_incoming=2, max_bounces=4, seed=3)
    media = scenes.media_box(width=16, height=16, scanline=5, time_bins=256, sqrt_spp=4, n_incoming=3,
                             max_bounces=12, seed=1)
    ico = scenes.icosphere_box(level=3, width=14, height=14, time_bins=300, sqrt_spp=2, n_incoming=2,
                               max_bounces=8, seed=2)
    rect = scenes.cornell_box(width=9, height=11, time_bins=64, sqrt_spp=2, n_incoming=1, max_bounces=5, seed=11)
    rect.rect_lights = np.array([[0.65, 0.99, 0.65, 0.65, 0.0, 0.65, 0.3, 0.3, 20.0]], dtype=np.float32)
    rect.film.camera_unwarp = True
    rect.film.offset = 0.5
    # a job in the Transient sampling mode beside steady-state ones
    media_ts = scenes.media_box(width=12, height=12, scanline=3, time_bins=200, sqrt_spp=3, n_incoming=2,
                                max_bounces=10, seed=21)
    media_ts.render.time_sampling = True
    return [box, media, ico, rect, media_ts]


def _single(R, s, **kw):
    img, st, _ = R.render_scene(s, **kw)
    return img, st


def _batch(R, scs, launch=None, autotune=False):
    b = R.Batch()
    setups = []
    for s in scs:
        w, cam, sampler, film, integ = R.setup_scene(s)
        b.add(w, integ, cam, film, sampler, seed=s.render.seed)
        setups.append((w, film))
    tuned = b.autotune() if autotune else None
    st = b.render(launch)
    return [f.read() for _, f in setups], st, tuned


def test_batch_equals_single_renders(R):
    scs = _jobs()
    imgs, st, _ = _batch(R, scs)
    ref = [_single(R, s) for s in scs]
    assert st["paths"] == sum(s.n_paths() for s in scs)
    assert st["extension_rays"] == sum(r[1]["extension_rays"] for r in ref)
    assert st["contributions"] == sum(r[1]["contributions"] for r in ref)
    for a, (b, _) in zip(imgs, ref):
        assert a.shape == b.shape
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-7)


@pytest.mark.parametrize("wave", [0, 1000, 4096])
def test_batch_wavefronts_split_jobs(R, wave):
    """Jobs straddle wavefront boundaries when the wavefront is small."""
    scs = _jobs()
    imgs, st, _ = _batch(R, scs, launch={"wave_paths": wave})
    ref = [_single(R, s)[0] for s in scs]
    if wave:
        assert st["waves"] == -(-sum(s.n_paths() for s in scs) // wave)
    for a, b in zip(imgs, ref):
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-7)


def test_graph_replay_equals_direct_launches(R):
    s = _jobs()[2]
    a, sa = _single(R, s, wave_paths=700)
    b, sb = _single(R, s, wave_paths=700, graph=True)
    assert sa["waves"] == sb["waves"] >= 3
    assert sa["kernel_launches"] == sb["kernel_launches"]
    assert sa["extension_rays"] == sb["extension_rays"]
    np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-7)
    imgs, st, _ = _batch(R, _jobs(), launch={"wave_paths": 2048, "flags": 8})  # TR_RENDER_GRAPH
    for x, sc in zip(imgs, _jobs()):
        np.testing.assert_allclose(x, _single(R, sc)[0], rtol=2e-5, atol=1e-7)


def test_per_job_mode_equals_single_renders(R):
    """TR_RENDER_PER_JOB renders the jobs of at least 2^22 paths one after another and the smaller
    ones together; the films and totals are unchanged."""
    from jarabotransientrendering_amd import _native as nat
    big = scenes.icosphere_box(level=2, width=64, height=64, time_bins=128, sqrt_spp=32, n_incoming=1,
                               max_bounces=4, seed=17)
    assert big.n_paths() >= 1 << 22
    scs = _jobs()[:2] + [big] + _jobs()[2:]
    imgs, st, _ = _batch(R, scs, launch={"flags": nat.TR_RENDER_PER_JOB})
    ref = [_single(R, s) for s in scs]
    assert st["paths"] == sum(s.n_paths() for s in scs)
    assert st["extension_rays"] == sum(r[1]["extension_rays"] for r in ref)
    for a, (b, _) in zip(imgs, ref):
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-7)


def test_autotune_then_render(R):
    scs = _jobs()
    imgs, st, (best, ms) = _batch(R, scs, autotune=True)
    assert best["wave_paths"] > 0 and best["trace_per_cu"] in (8, 16, 24, 32) and best["shade_per_cu"] in (1, 2, 4)
    assert ms > 0
    # the trials were cleared from the films: the render after tuning equals a plain render
    for a, sc in zip(imgs, scs):
        np.testing.assert_allclose(a, _single(R, sc)[0], rtol=2e-5, atol=1e-7)


def test_batch_errors(R):
    from jarabotransientrendering_amd import _native as nat
    b = R.Batch()
    with pytest.raises(nat.TrError):
        b.render()
    s = _jobs()[0]
    w, cam, sampler, film, integ = R.setup_scene(s)
    b.add(w, integ, cam, film, sampler)
    with pytest.raises(nat.TrError):
        b.add(w, integ, cam, film, sampler)  # same film twice
