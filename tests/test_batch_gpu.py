"""Batched solve of pairs with different sizes against single-pair solves.

The batch runs every pair's level through the same batched kernels (one
derivatives launch, the blocked primal-dual launches over all pairs' tiles, one
median/clamp launch per warp); each pair's flow must equal its own single-pair
blocked solve with the same tiles, bit for bit.
"""

import numpy as np
import pytest

from oracle import vfs_oracle as O
from conftest import flow_diff_stats

pytestmark = pytest.mark.gpu

SIZES = [(96, 128), (128, 128), (80, 112)]
TILES = {0: (16, 16, 4), 1: (16, 16, 4)}


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _params():
    import vfs_params as P
    prm = P.vfs_params()
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 10
    prm.Lambda, prm.alpha0, prm.alpha1, prm.beta, prm.gamma = 50.0, 5.0, 1.0, 10.0, 0.5
    return prm


def _pairs():
    import synthetic as S
    out = []
    for i, (h, w) in enumerate(SIZES):
        pr = S.make_pair(h, w, seed=100 + i, t=(0.2, 0.0, 0.02), rot_deg=1.5, rot_random=True, n_levels=2)
        cam = pr["cam"]
        tvf, cvf = O.get_vector_fields(O.get_grid_zero((h, w)), cam, cam, pr["R"], pr["t"], 1)
        out.append(dict(pair=pr, tvf=S.sanitize_field(tvf), cvf=cvf.astype(np.float32)))
    return out


def test_batch_equals_single_pair_solves(T):
    import copy
    import vfs as V
    import vfs_batch as B
    pairs = _pairs()
    prm = _params()
    batch = B.VFSBatch(prm, SIZES)
    batch.pd_tiles = TILES
    for i, c in enumerate(pairs):
        batch.setup_pair(i, c["pair"]["mask"], c["tvf"], c["cvf"])
    flows = batch.solve([c["pair"]["img_ref"] for c in pairs], [c["pair"]["img_other"] for c in pairs])
    flows = [f.cpu().numpy() for f in flows]
    again = [f.cpu().numpy() for f in batch.solve([c["pair"]["img_ref"] for c in pairs],
                                                  [c["pair"]["img_other"] for c in pairs])]
    for (h, w), c, f, f2 in zip(SIZES, pairs, flows, again):
        np.testing.assert_array_equal(f, f2)
        p = copy.copy(prm)
        p.height, p.width = h, w
        v = V.VFS(p)
        v.solver = "blocked"
        v.pd_tiles = TILES
        v.setup_module()
        v.compute_pyramid_sizes()
        v.create_mask_pyramid(c["pair"]["mask"])
        v.create_vector_field_pyramid(c["tvf"], c["cvf"])
        single = v.solve_stereo(c["pair"]["img_ref"], c["pair"]["img_other"]).cpu().numpy()
        np.testing.assert_array_equal(f, single)
        oprm = O.OracleParams(h, w, nLevel=2, fScale=2.0, nWarpIters=3, nSolverIters=10, Lambda=50.0, alpha0=5.0,
                              alpha1=1.0, beta=10.0, gamma=0.5)
        fo = O.solve_stereo(c["pair"]["img_ref"], c["pair"]["img_other"], c["pair"]["mask"], c["tvf"], c["cvf"], oprm)
        st = flow_diff_stats(f, fo, c["pair"]["mask"])
        print((h, w), "batched vs oracle:", st)
        assert st["median"] <= 1e-4


def test_batch_at_default_parameters_equals_single_solves(T):
    """The reference's default parameters and ratio (11 levels at fScale 1.2, lambda 5,
    alpha0 17, alpha1 1.2, beta 9, gamma 0.85; 2 warps x 8 iterations per level) on three
    pairs of uneven odd sizes: every level of every pair has its own size and the batched
    launches mix them; with the autotuned tiles of each launch's total pixel count, each
    flow equals its single-pair solve with the same tiles bit for bit."""
    import copy
    import synthetic as S
    import vfs as V
    import vfs_batch as B
    import vfs_params as P
    sizes = [(101, 133), (120, 136), (87, 97)]
    pairs = []
    for i, (h, w) in enumerate(sizes):
        pr = S.make_pair(h, w, seed=200 + i, t=(0.2, 0.0, 0.02), rot_deg=1.5, rot_random=True, n_levels=3,
                         level_ratio=1.2)
        tvf, cvf = O.get_vector_fields(O.get_grid_zero((h, w)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
        pairs.append(dict(pair=pr, tvf=S.sanitize_field(tvf), cvf=cvf.astype(np.float32)))
    prm = P.vfs_params()
    assert (prm.nLevel, prm.fScale) == (11, 1.2)
    prm.nWarpIters, prm.nSolverIters = 2, 8
    batch = B.VFSBatch(prm, sizes)
    for i, c in enumerate(pairs):
        batch.setup_pair(i, c["pair"]["mask"], c["tvf"], c["cvf"])
    flows = [f.cpu().numpy() for f in batch.solve([c["pair"]["img_ref"] for c in pairs],
                                                  [c["pair"]["img_other"] for c in pairs])]
    tiles = batch._tiles()
    for (h, w), c, f in zip(sizes, pairs, flows):
        p = copy.copy(prm)
        p.height, p.width = h, w
        v = V.VFS(p)
        v.solver = "blocked"
        v.pd_tiles = {lvl: tuple(t) for lvl, t in enumerate(tiles)}
        v.setup_module()
        v.compute_pyramid_sizes()
        assert len(v.sizes) == 11
        v.create_mask_pyramid(c["pair"]["mask"])
        v.create_vector_field_pyramid(c["tvf"], c["cvf"])
        single = v.solve_stereo(c["pair"]["img_ref"], c["pair"]["img_other"]).cpu().numpy()
        np.testing.assert_array_equal(f, single)
        assert np.isfinite(f).all() and np.abs(f).max() > 0


def test_batch_of_more_than_64_pairs_equals_single_solves(T):
    """70 small pairs in one batch: the launches' lookup of the pair that owns a tile or
    a pixel block scans the descriptors 64 at a time, so pairs 64-69 sit in the second
    group; pairs on both sides of that boundary equal their single-pair solves bit for
    bit."""
    import copy
    import synthetic as S
    import vfs as V
    import vfs_batch as B
    shapes = [(20, 24), (24, 20), (17, 23), (22, 22), (19, 26)]
    sizes = [shapes[i % len(shapes)] for i in range(70)]
    pairs = []
    for i, (h, w) in enumerate(sizes):
        pr = S.make_pair(h, w, seed=300 + i, t=(0.2, 0.0, 0.02), rot_deg=1.5, rot_random=True, n_levels=2)
        tvf, cvf = O.get_vector_fields(O.get_grid_zero((h, w)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
        pairs.append(dict(pair=pr, tvf=S.sanitize_field(tvf), cvf=cvf.astype(np.float32)))
    prm = _params()
    prm.nWarpIters, prm.nSolverIters = 2, 6
    batch = B.VFSBatch(prm, sizes)
    batch.pd_tiles = TILES
    for i, c in enumerate(pairs):
        batch.setup_pair(i, c["pair"]["mask"], c["tvf"], c["cvf"])
    flows = [f.cpu().numpy() for f in batch.solve([c["pair"]["img_ref"] for c in pairs],
                                                  [c["pair"]["img_other"] for c in pairs])]
    assert all(np.isfinite(f).all() for f in flows)
    for i in (0, 1, 62, 63, 64, 65, 68, 69):
        (h, w), c = sizes[i], pairs[i]
        p = copy.copy(prm)
        p.height, p.width = h, w
        v = V.VFS(p)
        v.solver = "blocked"
        v.pd_tiles = TILES
        v.setup_module()
        v.compute_pyramid_sizes()
        v.create_mask_pyramid(c["pair"]["mask"])
        v.create_vector_field_pyramid(c["tvf"], c["cvf"])
        single = v.solve_stereo(c["pair"]["img_ref"], c["pair"]["img_other"]).cpu().numpy()
        print(i, (h, w), "max |batch - single|", float(np.abs(flows[i] - single).max()))
        np.testing.assert_array_equal(flows[i], single)
        assert np.abs(flows[i]).max() > 0
