"""The C++ host driver (cli/render_scene.cpp over include/tr_bunnykiller.hpp and the C ABI):
it builds, parses the scene text format, fails loudly without a device, and on a GPU writes the
same streak slices as the Python host mirror for the same scene and seed."""
import json
import os
import subprocess

import numpy as np
import pytest

from jarabotransientrendering_amd import scenes
from jarabotransientrendering_amd.build import CLI


def _run(args, cwd):
    assert os.path.exists(CLI), "build the library and driver first (__graft_entry__.build())"
    return subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120)


def test_driver_usage_and_parse_errors(tmp_path):
    r = _run([], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr
    bad = tmp_path / "bad.txt"
    bad.write_text("film 4 4 8 0.01 0 -1 0\nsphere 0 0 0 1\n")
    r = _run([str(bad)], tmp_path)
    assert r.returncode == 1 and "unknown directive sphere" in r.stderr
    bad.write_text("material 1 1 1\ntri 0 0 0 1 0 0 0 1 0 3\n")
    r = _run([str(bad)], tmp_path)
    assert r.returncode == 1 and "bad material index" in r.stderr


def test_driver_fails_loudly_without_device(tmp_path):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    s = scenes.cornell_box(width=4, height=4, time_bins=8, sqrt_spp=1, n_incoming=1, max_bounces=2, seed=1)
    (tmp_path / "s.txt").write_text(s.to_ref_text("out"))
    r = _run(["s.txt"], tmp_path)
    assert r.returncode == 1 and "device" in r.stderr.lower()
    assert not list(tmp_path.glob("out_*.raw"))


def _slices(prefix, rows, W, T):
    return np.stack([np.fromfile(f"{prefix}_{y:04d}.raw", dtype=np.float32).reshape(W, T, 3) for y in rows])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["box", "media_scanline"])
def test_driver_matches_python_host(tmp_path, kind):
    from jarabotransientrendering_amd import renderer
    if kind == "box":
        s = scenes.cornell_box(width=12, height=10, time_bins=128, sqrt_spp=2, n_incoming=3, max_bounces=5, seed=7)
        s.rect_lights = np.array([[0.65, 0.99, 0.65, 0.65, 0.0, 0.65, 0.3, 0.3, 20.0]], dtype=np.float32)
        s.film.offset = 0.25
        s.film.camera_unwarp = True
    else:
        s = scenes.media_box(width=16, height=16, scanline=5, time_bins=256, sqrt_spp=2, n_incoming=2,
                             max_bounces=8, seed=5)
    (tmp_path / "s.txt").write_text(s.to_ref_text("out"))
    r = _run(["s.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout.strip().splitlines()[-1])
    img, pst, _ = renderer.render_scene(s)
    assert st["paths"] == pst["paths"] == s.n_paths()
    assert st["extension_rays"] == pst["extension_rays"] and st["contributions"] == pst["contributions"]
    rows = [s.film.scanline] if s.film.scanline >= 0 else list(range(s.film.height))
    raw = _slices(str(tmp_path / "out"), rows, s.film.width, s.film.time_bins)
    # same random streams and geometry order; only float-atomic summation order differs
    np.testing.assert_allclose(raw, img, rtol=1e-5, atol=1e-7)
    assert "Total time" in (tmp_path / "out.log").read_text()
