"""The tuned tile table (vfs-python_amd/autotune_gfx950.json) against the compiled
variants, and the lookups the solvers use (CPU; the library is only asked for its
variant list)."""

import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "vfs-python_amd"))


def _doc():
    with open(os.path.join(ROOT, "vfs-python_amd", "autotune_gfx950.json")) as f:
        return json.load(f)


def _variants():
    import vfs_hip as H
    try:
        return {tuple(v) for v in H.pd_variants()}
    except (RuntimeError, OSError) as e:
        pytest.skip(f"library not built: {e}")


def test_every_table_entry_is_a_compiled_variant():
    compiled = _variants()
    doc = _doc()
    for e in doc["entries"] + doc.get("batch_entries", []):
        assert tuple(e["best"]) in compiled, e


def test_single_level_lookup_prefers_the_exact_entry_then_the_nearest():
    import autotune as A
    doc = _doc()
    for e in doc["entries"]:
        assert A.tile_for(e["h"], e["w"], e["iters"]) == tuple(e["best"])
    # a size between two entries takes the one nearest in pixel count
    sizes = sorted({(e["h"] * e["w"], e["h"], e["w"]) for e in doc["entries"] if e["iters"] == 50})
    (n0, h0, w0), (n1, h1, w1) = sizes[-2], sizes[-1]
    mid = int(((n0 + 0.1 * (n1 - n0)) ** 0.5))
    want = next(tuple(e["best"]) for e in doc["entries"] if (e["h"], e["w"], e["iters"]) == (h0, w0, 50))
    assert A.tile_for(mid, mid, 50) == want


def test_batch_lookup_uses_batch_entries_within_their_range():
    import autotune as A
    doc = _doc()
    be = doc["batch_entries"]
    for e in be:
        assert A.tile_for_pixels(e["pixels"], e["iters"]) == tuple(e["best"])
    # below half the smallest batch entry the single-level entry nearest in pixel count decides
    small = min(e["pixels"] for e in be) // 4
    import math
    near = min((e for e in doc["entries"] if e["iters"] == 50),
               key=lambda e: abs(math.log(e["h"] * e["w"]) - math.log(small)))
    assert A.tile_for_pixels(small, 50) == tuple(near["best"])
