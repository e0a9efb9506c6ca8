"""Host-side helpers of bench.py: the extended-tile pixel count behind every algorithmic
byte figure, the counter table lookup of the gradient leg, and the --gpus guard."""

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _ext_px_brute(hh, ww, tx, ty, k):
    """Every tile's extended region painted pixel by pixel (clipped to the image)."""
    tot = 0
    for y0 in range(0, hh, ty):
        for x0 in range(0, ww, tx):
            cover = np.zeros((hh, ww), bool)
            cover[max(0, y0 - k):min(hh, y0 + ty + k), max(0, x0 - k):min(ww, x0 + tx + k)] = True
            tot += int(cover.sum())
    return tot


@pytest.mark.parametrize("hh,ww,tx,ty,k", [(50, 50, 12, 12, 10), (97, 131, 22, 22, 5), (8, 9, 16, 16, 8),
                                           (100, 68, 26, 10, 3), (33, 17, 20, 20, 6)])
def test_extended_pixels_match_brute_force(hh, ww, tx, ty, k):
    assert bench._ext_px(hh, ww, tx, ty, k) == _ext_px_brute(hh, ww, tx, ty, k)


def test_extended_pixels_closed_form_for_a_whole_tile_grid():
    # 800 = 32 tiles of 25: each spans 25 + 2 * 3 pixels per axis, the first and last clipped by 3 on
    # their outer side; the count is the product of the per-axis sums
    per_axis = 32 * 31 - 2 * 3
    assert bench._ext_px(800, 800, 25, 25, 3) == per_axis * per_axis


def test_adjoint_counter_lookup_reads_the_committed_table():
    t = bench.measured_adj_traffic([26, 10, 3, 512])
    assert t is not None and 50e6 < t < 1e9
    assert bench.measured_adj_traffic([7, 7, 7, 7]) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
