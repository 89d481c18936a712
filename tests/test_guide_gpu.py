"""GPU checks of the guiding network (tr_guide_* through the C ABI) against the CPU restatement
(oracle/guide_ref.py), and of guided rendering: the one-sample MIS estimator must stay unbiased
whatever the network outputs."""
import numpy as np
import pytest

from jarabotransientrendering_amd import scenes
from oracle import guide_ref as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    from jarabotransientrendering_amd import renderer
    return renderer


def _raw(rng, n):
    raw = np.zeros((n, 12), np.float32)
    raw[:, 0:3] = rng.random((n, 3))
    wo = rng.normal(size=(n, 3))
    raw[:, 3:6] = wo / np.linalg.norm(wo, axis=1, keepdims=True)
    med = rng.random(n) < 0.3
    nn = rng.normal(size=(n, 3))
    raw[:, 6:9] = np.where(med[:, None], 0, nn / np.linalg.norm(nn, axis=1, keepdims=True))
    raw[:, 9] = rng.random(n)
    raw[:, 10] = med
    return raw


def _records(rng, n):
    raw = _raw(rng, n)
    rec = np.zeros((n, 16), np.float32)
    rec[:, 0:3] = raw[:, 0:3]
    rec[:, 3] = raw[:, 9]
    rec[:, 4:7] = raw[:, 3:6]
    rec[:, 7] = raw[:, 10]
    rec[:, 8:11] = raw[:, 6:9]
    rec[:, 11] = rng.exponential(1.0, n) * (rng.random(n) < 0.8)  # some zero-weight records
    om = rng.normal(size=(n, 3))
    rec[:, 12:15] = om / np.linalg.norm(om, axis=1, keepdims=True)
    return rec


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4099])
def test_eval_matches_cpu_restatement(R, n):
    """Matrix-core forward pass (BF16 operands, FP32 accumulation) vs the float64 restatement with
    the same BF16 roundings; also pins the MFMA lane layouts and the K permutation."""
    g = R.PathGuide(init_seed=3)
    p = G.initial_like(seed=7, out_scale=1.0)
    g.params = p
    assert np.array_equal(g.params, p)
    rng = np.random.default_rng(n)
    raw = _raw(rng, n)
    out = g.eval(raw)
    ref = G.forward(p, raw, emulate_bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(out - ref).max()
    # FP32 vs float64 accumulation can flip a bf16 rounding of a hidden activation: ~2^-8 relative
    assert err < 1.5e-2 * scale, (err, scale)
    assert np.abs(out - ref).mean() < 1e-3 * scale


def test_gradient_matches_cpu_restatement(R):
    g = R.PathGuide(init_seed=3, batch_size=1024)
    p = G.initial_like(seed=8, out_scale=1.0)
    g.params = p
    rng = np.random.default_rng(9)
    rec = _records(rng, 700)
    grad, loss = g.gradient(rec)
    ref = G.gradient(p.astype(np.float64), rec.astype(np.float64), emulate_bf16=True)
    exact = G.gradient(p.astype(np.float64), rec.astype(np.float64))
    ref_loss = G.loss(p, rec, emulate_bf16=True)
    assert abs(loss - ref_loss) < 2e-2 * abs(ref_loss) + 1e-4
    # against the same BF16 roundings: only FP32-vs-float64 accumulation differs
    rel = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    cos = grad @ ref / np.linalg.norm(grad) / np.linalg.norm(ref)
    assert rel < 2e-2 and cos > 0.9998, (rel, cos)
    # against the exact float64 gradient: BF16 training is a few percent off, same direction
    rel_exact = np.linalg.norm(grad - exact) / np.linalg.norm(exact)
    assert rel_exact < 0.1 and grad @ exact / np.linalg.norm(grad) / np.linalg.norm(exact) > 0.995
    # every layer's block (weights and biases) is produced
    k = 0
    for i, o in G.LAYERS:
        blk = slice(k, k + i * o + o)
        k += i * o + o
        lr = np.linalg.norm(grad[blk] - ref[blk]) / max(np.linalg.norm(ref[blk]), 1e-12)
        assert lr < 0.1, (i, o, lr)


def test_adam_step_matches_textbook_update(R):
    g = R.PathGuide(init_seed=3, batch_size=1024, learning_rate=1e-3)
    p = G.initial_like(seed=10, out_scale=1.0)
    g.params = p
    rec = _records(np.random.default_rng(11), 512)
    grad, _ = g.gradient(rec)
    g.train_batch(rec)
    p1 = g.params
    expect, _, _ = G.adam(p.astype(np.float64), 0.0, 0.0, grad.astype(np.float64), t=1, lr=1e-3)
    np.testing.assert_allclose(p1, expect, rtol=0, atol=2e-6)
    assert g.info()["steps"] == 1


def test_training_reduces_loss_on_a_fixed_batch(R):
    g = R.PathGuide(init_seed=5, batch_size=2048, learning_rate=3e-3)
    rng = np.random.default_rng(12)
    rec = _records(rng, 2048)
    # a learnable target: directions cluster around a function of position
    d = np.stack([np.sin(6 * rec[:, 0]), np.cos(6 * rec[:, 1]), rec[:, 2] - 0.5], axis=1)
    d += 0.1 * rng.normal(size=d.shape)
    rec[:, 12:15] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 11] = 1.0
    losses = [g.train_batch(rec) for _ in range(150)]
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < np.mean(losses[:10]) - 0.3, (losses[:3], losses[-3:])


def test_errors(R):
    from jarabotransientrendering_amd import _native as nat
    with pytest.raises(nat.TrError):
        R.PathGuide(batch_size=100)
    with pytest.raises(nat.TrError):
        R.PathGuide(bsdf_fraction=0.0)
    g = R.PathGuide(batch_size=64)
    with pytest.raises(nat.TrError):
        g.train_batch(np.zeros((65, 16), np.float32))
    bad = np.zeros(nat.GUIDE_PARAMS, np.float32)
    bad[5] = np.nan
    with pytest.raises(nat.TrError):
        g.params = bad
    # guiding does not combine with the Transient sampling mode
    s = scenes.media_box(width=4, height=4, scanline=1, time_bins=16, sqrt_spp=1, n_incoming=1, max_bounces=3)
    s.render.time_sampling = True
    with pytest.raises(nat.TrError, match="TIME_SAMPLING"):
        R.render_scene(s, guide=g)


# ---------------------------------------------------------------------------------------------
# guided rendering

def _per_pixel_sums(img):
    return img.sum(axis=(2, 3))


def _mean_and_se(R, scene, seeds, **kw):
    imgs = [_per_pixel_sums(R.render_scene(scene, seed=s, **kw)[0]) for s in seeds]
    a = np.stack(imgs)
    return a.mean(0), a.std(0, ddof=1) / np.sqrt(len(seeds)), a


@pytest.mark.parametrize("kind,fraction", [("box", 0.5), ("media", 0.5), ("media", 0.1)])
def test_guided_render_is_unbiased(R, kind, fraction):
    """Guided (network mixture + BSDF, one-sample MIS) and plain renders estimate the same film.
    The network is deliberately mis-trained (sharp lobes in fixed directions) so that a wrong
    density in the MIS weight would show as a bias."""
    if kind == "box":
        s = scenes.cornell_box(width=8, height=8, time_bins=32, exposure=0.1, sqrt_spp=8, n_incoming=4,
                               max_bounces=5, seed=0)
    else:
        s = scenes.media_box(width=8, height=8, scanline=-1, time_bins=32, exposure=0.1, sqrt_spp=8,
                             n_incoming=4, max_bounces=8, seed=0)
    p = G.initial_like(seed=13, out_scale=1.0)
    k = G.N_PARAMS - 40
    p[k:] = 0.0
    p[k + 3::5] = 4.0                       # kappa ~ 4 (sharp-ish lobes)
    p[k + 0:k + 40:5] = np.linspace(-1, 1, 8)  # lobe directions spread by the output bias
    p[k + 1:k + 40:5] = 1.0
    guide = R.PathGuide(bsdf_fraction=fraction, steps_per_wave=0, init_seed=1)
    guide.params = p
    seeds = list(range(12))
    mg, sg, _ = _mean_and_se(R, s, seeds, guide=guide, wave_paths=4096)
    mp, sp, _ = _mean_and_se(R, s, [100 + x for x in seeds], orthonormal_frame=True)
    z = (mg - mp) / np.sqrt(sg ** 2 + sp ** 2 + 1e-30)
    # per-pixel time-integrated radiance: no more than a few pixels beyond 4 sigma, and the image
    # totals agree within their combined standard error
    assert (np.abs(z) > 4).sum() <= 1, np.sort(np.abs(z).ravel())[-5:]
    tot_se = np.sqrt((sg ** 2).sum() + (sp ** 2).sum())
    assert abs(mg.sum() - mp.sum()) < 4 * tot_se, (mg.sum(), mp.sum(), tot_se)


def test_orthonormal_frame_matches_reference_frame_on_axis_aligned_geometry(R):
    """On the box (normals along the axes) the reference's shading frame is orthonormal too, so
    both samplers estimate the same film."""
    s = scenes.cornell_box(width=8, height=8, time_bins=32, exposure=0.1, sqrt_spp=8, n_incoming=4, max_bounces=5,
                           seed=0)
    seeds = list(range(10))
    ma, sa, _ = _mean_and_se(R, s, seeds)
    mb, sb, _ = _mean_and_se(R, s, [50 + x for x in seeds], orthonormal_frame=True)
    z = (ma - mb) / np.sqrt(sa ** 2 + sb ** 2 + 1e-30)
    assert (np.abs(z) > 4).sum() <= 1
    assert abs(ma.sum() - mb.sum()) < 4 * np.sqrt((sa ** 2).sum() + (sb ** 2).sum())


def test_guided_render_trains_online(R):
    s = scenes.media_box(width=16, height=16, scanline=-1, time_bins=64, exposure=0.05, sqrt_spp=4, n_incoming=4,
                         max_bounces=8, seed=0)
    guide = R.PathGuide(train_paths=2048, batch_size=2048, steps_per_wave=1)
    img, st, _ = R.render_scene(s, guide=guide, wave_paths=1 << 13, profile=True)
    info = guide.info()
    assert info["steps"] == st["waves"] and st["waves"] == s.n_paths() // (1 << 13)
    assert np.isfinite(info["last_loss"])
    assert np.isfinite(img).all() and img.sum() > 0
    assert st["kernel_launches"]["guide"] == 2 * st["waves"]
    # frozen: later renders take no Adam step and keep the weights
    guide.set_steps(0)
    before = guide.params
    img2, st2, _ = R.render_scene(s, guide=guide, wave_paths=1 << 13, seed=5)
    assert guide.info()["steps"] == info["steps"]
    np.testing.assert_array_equal(guide.params, before)
    assert np.isfinite(img2).all() and img2.sum() > 0
