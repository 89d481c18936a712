"""CPU check of the multibounce wavefront clamp (csrc/tr_internal.h, mb_wave): a multibounce
shadow record tags (camera sample within the wavefront) << 5 | layer in 32 bits, so a wavefront
must hold whole camera samples and fewer than 2^27 of them, for any requested wavefront size and
paths per camera sample. Compiled with g++ from the header itself."""
import os
import subprocess

from tests.conftest import ROOT

PROG = r"""
#include <cstdio>
#include "tr_internal.h"
int main() {
    const uint32_t waves[] = {1u, 2u, 3u, 1000u, 1u << 25, (1u << 27) - 1u, 1u << 27, (1u << 27) + 5u,
                              1u << 30, 0xFFFFFFFFu};
    const uint32_t incs[] = {1u, 2u, 3u, 5u, 7u, 64u, 1000u};
    for (uint32_t w : waves)
        for (uint32_t n : incs) std::printf("%u %u %u\n", w, n, mb_wave(w, n));
    return 0;
}
"""


def test_mb_wave_keeps_camera_sample_tags_in_27_bits(tmp_path):
    src = tmp_path / "mbw.cpp"
    src.write_text(PROG)
    exe = tmp_path / "mbw"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "jarabotransientrendering_amd", "csrc"),
                           str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    rows = [tuple(int(v) for v in line.split()) for line in out if line.strip()]
    assert len(rows) == 70
    for w, n, got in rows:
        assert got % n == 0 and got >= n                      # whole camera samples, at least one
        assert got // n < (1 << 27)                            # camera-sample index fits the tag
        assert (((got // n - 1) << 5) | 31) < (1 << 32)        # largest tag fits 32 bits
        if w >= n and w // n < (1 << 27):
            assert got == w // n * n                           # below the cap: the request, rounded down
