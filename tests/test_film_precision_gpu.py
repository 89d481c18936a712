"""Film accumulation: compensated per-bin sums on the GPU, pinned against exact (FP64 / fsum)
totals, plus the reference's binning rule and its non-finite-sample rule.

StreakCameraFilm::add_samples adds each sample to its bin one after another in FP32
(StreakCameraFilm.h:139-192, 280-297); a bin that receives many small samples after a large one
loses them in a plain FP32 running sum. The film here keeps (sum, compensation) pairs and makes
every addition (LDS partials, HBM flushes, direct deposits) compensated, so the total is exact up
to second-order terms."""
import math

import numpy as np
import pytest

from jarabotransientrendering_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    from jarabotransientrendering_amd import renderer
    return renderer


def _samples(n):
    from jarabotransientrendering_amd import _native as nat
    return np.zeros(n, dtype=nat.FILM_SAMPLE_DTYPE)


def _fsum32(v):
    return math.fsum(np.asarray(v, dtype=np.float64).tolist())


@pytest.mark.parametrize("pixels", [1, 16])
def test_one_bin_many_small_samples_is_exact(R, pixels):
    """One large sample, then 2^22 samples below half an ulp of it, in one bin per pixel. A plain
    FP32 running sum keeps only the large one (its error bound is n * u * sum|x|, u = 2^-24); the
    GPU film's sum + compensation stays within the compensated-summation bound
    (2u + O(n u^2)) * sum|x| (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed.,
    sec. 4.3; the O(n u^2) term taken as 8 n u^2), whatever the order of the additions, and at least
    10^4 times closer to the exact total than the plain sum. pixels=1 keeps every record in the LDS
    histogram of its block; 16 pixels overflow the two LDS pixel slots a block has at 1024 bins, so
    most records go straight to HBM, one compensated atomic addition each."""
    n = 1 << 22
    rng = np.random.default_rng(7)
    small = rng.uniform(2.0 ** -26, 2.0 ** -25, n).astype(np.float32)
    s = _samples(n + pixels)
    s["x"][:pixels] = np.arange(pixels)
    s["rgb"][:pixels] = 1.0
    s["x"][pixels:] = np.arange(n) % pixels
    s["rgb"][pixels:, 0] = small
    s["rgb"][pixels:, 1] = small * np.float32(0.5)
    s["rgb"][pixels:, 2] = small * np.float32(3.0)
    s["time"] = np.float32(0.0535)   # bin 5 of 0.01
    s["weight"][:pixels] = 1.0
    film = R.StreakCameraFilm(pixels, 1, 1024, 0.01)
    film.add_samples(s)
    sums, w = film.read_sums()
    assert sums.shape == (1, pixels, 1024, 3)
    np.testing.assert_array_equal(w, np.ones((1, pixels)))
    nz = np.argwhere(sums != 0)
    assert set(map(int, nz[:, 2])) == {5}
    for p in range(pixels):
        for c in range(3):
            vals = s["rgb"][s["x"] == p][:, c]
            exact = _fsum32(vals)
            naive = np.cumsum(vals, dtype=np.float32)[-1]
            # the failure mode being guarded against: a plain FP32 running sum loses the small samples
            assert abs(naive - exact) > 2e-3 * exact
            got = sums[0, p, 5, c]
            u = 2.0 ** -24
            bound = (2 * u + 8 * len(vals) * u * u) * float(np.abs(vals.astype(np.float64)).sum())
            assert abs(got - exact) <= bound, (p, c, got, exact, naive, bound)
            assert abs(got - exact) * 1e4 <= abs(float(naive) - exact), (p, c, got, exact, naive)


def test_add_samples_bins_like_the_reference(R):
    """bin = floor((t - offset) / exposure) in FP32, kept if 0 <= pos < T (StreakCameraFilm.h:165-171);
    the weight counts for samples outside the window too."""
    T, exposure, offset = 64, 0.02, 0.1
    rng = np.random.default_rng(3)
    n = 5000
    s = _samples(n)
    s["x"] = rng.integers(0, 4, n)
    s["y"] = rng.integers(0, 3, n)
    s["time"] = rng.uniform(0.0, offset + T * exposure * 1.2, n).astype(np.float32)
    s["time"][:4] = np.float32(offset), np.float32(offset + T * exposure), np.float32(offset - 1e-6), np.float32(0.5)
    s["rgb"] = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    s["weight"] = rng.integers(0, 2, n).astype(np.float32)
    film = R.StreakCameraFilm(4, 3, T, exposure)
    film.set_offset(offset)
    film.add_samples(s)
    sums, w = film.read_sums()
    pos = (s["time"] - np.float32(offset)) / np.float32(exposure)
    keep = (pos < T) & (pos >= 0)
    ref = np.zeros((3, 4, T, 3))
    bins = np.minimum(np.floor(pos[keep]).astype(np.int64), T - 1)
    np.add.at(ref, (s["y"][keep], s["x"][keep], bins), s["rgb"][keep].astype(np.float64))
    np.testing.assert_allclose(sums, ref, rtol=1e-6, atol=1e-6)
    wref = np.zeros((3, 4))
    np.add.at(wref, (s["y"], s["x"]), s["weight"].astype(np.float64))
    np.testing.assert_array_equal(w, wref)


def test_add_samples_rejects_samples_off_the_film(R):
    from jarabotransientrendering_amd import _native as nat
    s = _samples(2)
    s["x"] = [0, 4]
    film = R.StreakCameraFilm(4, 2, 16, 0.1)
    with pytest.raises(nat.TrError):
        film.add_samples(s)
    sums, w = film.read_sums()
    assert not sums.any() and not w.any()


def test_render_accumulates_passes_compensated(R):
    """A render adds into the film like add_samples: a render after a large caller sample in the
    same bin keeps every path's contribution (the render alone, added to the large value in FP64,
    gives the same totals)."""
    s = scenes.cornell_box(width=4, height=4, time_bins=128, exposure=0.02, sqrt_spp=8, max_bounces=3, seed=2)
    films = []
    for big in (False, True):
        w, cam, sampler, film, integ = R.setup_scene(s)
        if big:
            q = _samples(4 * 4 * 128)
            q["x"] = np.tile(np.repeat(np.arange(4), 128), 4)
            q["y"] = np.repeat(np.arange(4), 4 * 128)
            q["time"] = np.tile((np.arange(128) + 0.5) * 0.02, 16).astype(np.float32)
            q["rgb"] = 4096.0
            film.add_samples(q)
        R.TransientRenderer().render(None, w, integ, cam, film, sampler, seed=s.render.seed)
        films.append(film.read_sums()[0])
    alone, with_big = films
    # the bins hold 4096 + (render's value); the render's part comes back within FP32 rounding of it
    np.testing.assert_allclose(with_big - 4096.0, alone, rtol=1e-5, atol=4096.0 * 2.0 ** -40)


def test_non_finite_samples_are_zeroed(R, oracle):
    """TransientRenderer.h:167-171: non-finite sample radiance becomes 0 (the camera sample's weight
    still counts). A light of infinite intensity therefore adds nothing: the film equals the film
    rendered with that light at zero intensity on the same random streams, and the oracle's."""
    base = scenes.cornell_box(width=8, height=8, time_bins=64, exposure=0.04, sqrt_spp=2, max_bounces=3, seed=9)
    films = {}
    for label, val in (("inf", np.inf), ("nan", np.nan), ("zero", 0.0)):
        s = scenes.cornell_box(width=8, height=8, time_bins=64, exposure=0.04, sqrt_spp=2, max_bounces=3, seed=9)
        # a non-finite red channel makes the whole sample non-finite; "zero" is a black light
        extra = np.array([[0.3, 0.9, 0.4, val, val != 0, val != 0]], dtype=np.float32)
        s.point_lights = np.concatenate([base.point_lights, extra]).astype(np.float32)
        img, stats, _ = R.render_scene(s)
        assert np.isfinite(img).all()
        films[label] = (img, s)
    np.testing.assert_array_equal(films["inf"][0], films["zero"][0])
    np.testing.assert_array_equal(films["nan"][0], films["zero"][0])
    assert films["zero"][0].sum() > 0
    f, wgt, _ = oracle.render(films["inf"][1], oracle.RNG_PER_PATH)
    ref = oracle.normalize(f, wgt)
    assert np.isfinite(ref).all()
    np.testing.assert_allclose(films["inf"][0].sum(), ref.sum(), rtol=1e-3)
