"""Image output (tr_image_save / tr_film_write): Radiance RGBE files are byte for byte the files
the reference's Imaging::save wrote for the same float images (tests/golden/image_*.npz), raw dumps
are the float data, and the decoder used here is pinned by the reference's own bytes."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN


def decode_rgbe(data: bytes):
    """Radiance picture -> (RGBE bytes [H][W][4], floats [H][W][3])."""
    end = data.index(b"\n\n")
    header = data[:end].decode()
    assert header.startswith("#?RADIANCE") and "FORMAT=32-bit_rle_rgbe" in header
    nl = data.index(b"\n", end + 2)
    res = data[end + 2:nl].decode().split()
    assert res[0] == "-Y" and res[2] == "+X"
    H, W = int(res[1]), int(res[3])
    p = nl + 1
    out = np.zeros((H, W, 4), np.uint8)
    for y in range(H):
        if 8 <= W <= 32767 and data[p] == 2 and data[p + 1] == 2 and data[p + 2] < 128:
            assert (data[p + 2] << 8 | data[p + 3]) == W
            p += 4
            for c in range(4):
                x = 0
                while x < W:
                    n = data[p]
                    p += 1
                    if n > 128:
                        out[y, x:x + n - 128, c] = data[p]
                        p += 1
                        x += n - 128
                    else:
                        assert n > 0
                        out[y, x:x + n, c] = np.frombuffer(data[p:p + n], np.uint8)
                        p += n
                        x += n
        else:
            out[y] = np.frombuffer(data[p:p + 4 * W], np.uint8).reshape(W, 4)
            p += 4 * W
    assert p == len(data)
    e = out[..., 3].astype(np.int32)
    f = np.where(e[..., None] > 0, (out[..., :3].astype(np.float64)) * np.ldexp(1.0, e - 136)[..., None], 0.0)
    return out, f


@pytest.mark.parametrize("tag", ["narrow", "runs", "streak_like"])
def test_rgbe_matches_reference_imaging_save(tmp_path, tag):
    from jarabotransientrendering_amd import _native as nat
    g = np.load(os.path.join(GOLDEN, f"image_{tag}.npz"))
    img = np.ascontiguousarray(g["image"], dtype=np.float32)
    H, W, _ = img.shape
    ref_q, ref_f = decode_rgbe(g["hdr"].tobytes())
    path = str(tmp_path / "out.hdr")
    assert nat.lib().tr_image_save(path.encode(), nat.as_ptr(img), W, H) == 0
    mine = open(path, "rb").read()
    assert mine == g["hdr"].tobytes()          # the reference's file, byte for byte
    mine_q, mine_f = decode_rgbe(mine)
    assert mine_q.shape == (H, W, 4)
    np.testing.assert_array_equal(mine_q, ref_q)
    # and the decoded values are the (non-negative) inputs to RGBE precision: 8 mantissa bits of
    # the largest channel
    ok = (img >= 0).all(axis=2)
    m = img.max(axis=2, keepdims=True)
    assert np.all((np.abs(mine_f - img) <= m / 128 + 1e-30)[ok])


def test_raw_and_errors(tmp_path):
    from jarabotransientrendering_amd import _native as nat
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    p = str(tmp_path / "a.raw")
    assert nat.lib().tr_image_save(p.encode(), nat.as_ptr(img), 3, 2) == 0
    assert np.array_equal(np.fromfile(p, np.float32), img.ravel())
    assert nat.lib().tr_image_save(str(tmp_path / "a.png").encode(), nat.as_ptr(img), 3, 2) == nat.TR_ERR_INVALID
    assert b"unsupported" in nat.lib().tr_last_error()
    assert nat.lib().tr_image_save(b"/nonexistent_dir/x.hdr", nat.as_ptr(img), 3, 2) == nat.TR_ERR_INVALID


@pytest.mark.gpu
@pytest.mark.parametrize("T,scanline", [(64, -1), (5, 3)])
def test_film_write_files(tmp_path, T, scanline):
    """Streak slices (transposed below 8 time bins), raw and RGBE, and the steady image."""
    from jarabotransientrendering_amd import renderer, scenes
    s = scenes.cornell_box(width=10, height=6, time_bins=T, exposure=2.56 / T, sqrt_spp=2, n_incoming=2,
                           max_bounces=3, seed=4)
    s.film.scanline = scanline
    img, st, film = renderer.render_scene(s)
    rows = [scanline] if scanline >= 0 else list(range(6))
    paths = film.write(str(tmp_path / "raw" / "f"), "raw")
    assert paths[-1].endswith("f.hdr")
    for i, y in enumerate(rows):
        a = np.fromfile(paths[i], np.float32)
        slice_ = img[i] if T >= 8 else img[i].transpose(1, 0, 2)
        assert np.array_equal(a, np.ascontiguousarray(slice_).ravel())
    hdr_paths = film.write(str(tmp_path / "f"), "hdr")
    for i, y in enumerate(rows):
        q, f = decode_rgbe(open(hdr_paths[i], "rb").read())
        slice_ = img[i] if T >= 8 else img[i].transpose(1, 0, 2)
        assert f.shape == slice_.shape
        m = slice_.max(axis=2, keepdims=True)
        assert np.all(np.abs(f - slice_) <= m / 128 + 1e-30)
    q, steady = decode_rgbe(open(str(tmp_path / "f.hdr"), "rb").read())
    assert steady.shape == (6, 10, 3)
    expect = np.zeros((6, 10, 3))
    for i, y in enumerate(rows):
        expect[y] = img[i].sum(axis=1)
    assert np.all(np.abs(steady - expect) <= expect.max(axis=2, keepdims=True) / 128 + 1e-30)
