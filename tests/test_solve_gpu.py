"""End-to-end parity of VFS.solve_stereo / triangulate_flow on the GPU against the oracle.

Workload: the seeded 256x256 synthetic fisheye pair (f = 256/pi, t = (0.2, 0, 0) m,
1 deg rotation), 3 levels x 10 warps x 20 primal-dual iterations, lambda = 50,
alpha0 = 5, alpha1 = 1, beta = 10, gamma = 0.5.

Tolerances (fp32 GPU against the float64 oracle after the full iteration count):
  * coarsest level (64x64, after its 10 warps): median |d flow| <= 1e-6 px,
    99th percentile <= 1e-3 px;
  * full resolution: median |d flow| <= 1e-4 px (the north star's figure holds as a
    median); 90th percentile <= 5e-2 px and 99th percentile <= 0.5 px.  Measured:
    median 3.6e-5, p90 1.4e-2, p99 0.20, max 12.1 px.  The tail is the iteration's
    amplification of rounding, not a kernel error: this workload's lambda = 50 makes
    the solve diverge locally (a few pixels run off by many pixels in any precision),
    and the float32 NumPy restatement of the same formulas deviates from the float64
    oracle with the same distribution (median 3.7e-5, p90 1.3e-2, p99 0.20, max
    9.4 px), so no per-pixel bound like max |d flow| <= 1e-4 px is meaningful here;
    the GPU is also held to <= 3x that restatement's p99;
  * depth: median relative depth difference <= 1e-3 over the mask;
  * the final mask (flow zero outside it) is exact.
"""

import numpy as np
import pytest

from oracle import vfs_oracle as O
from conftest import flow_diff_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solver(small_case, mode):
    import vfs as V
    import vfs_params as P
    import synthetic as S
    pair = small_case["pair"]
    prm = P.vfs_params(pair["img_ref"])
    S.apply_workload_params(prm, "plain_small_256")
    v = V.VFS(prm)
    v.solver = mode
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pair["mask"])
    v.create_vector_field_pyramid(small_case["tvf"], small_case["cvf"])
    return v


def test_solve_graph_matches_oracle(T, small_case, small_oracle):
    v = _solver(small_case, "graph")
    flow = v.solve_stereo(small_case["pair"]["img_ref"], small_case["pair"]["img_other"])
    T.cuda.synchronize()
    f = flow.cpu().numpy()
    mask = small_case["pair"]["mask"]
    # coarsest level state after its warps (the level buffers of the graph solve)
    coarse = v._bufs["levels"][-1]["flow"].cpu().numpy()
    mk = v.masks[-1].cpu().numpy()
    c = flow_diff_stats(coarse, small_oracle["trace"][2]["warp"], mk)
    st = flow_diff_stats(f, small_oracle["flow"], mask)
    st32 = flow_diff_stats(small_oracle["flow32"], small_oracle["flow"], mask)
    print("coarsest level vs oracle:", c)
    print("full resolution vs oracle:", st)
    print("fp32 restatement vs oracle:", st32)
    assert np.all(f[mask == 0] == 0)
    assert c["median"] <= 1e-6 and c["p99"] <= 1e-3
    assert st["median"] <= 1e-4
    assert st["p90"] <= 5e-2 and st["p99"] <= 0.5
    assert st["p99"] <= 3 * st32["p99"]


def test_solver_modes_agree_bitwise(T, small_case):
    pair = small_case["pair"]
    outs = {}
    for mode in ("graph", "eager", "python"):
        v = _solver(small_case, mode)
        outs[mode] = v.solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
        if mode == "graph":  # replay of the recorded graph gives the same result
            again = v.solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
            np.testing.assert_array_equal(again, outs[mode])
    np.testing.assert_array_equal(outs["graph"], outs["eager"])
    np.testing.assert_array_equal(outs["graph"], outs["python"])


@pytest.mark.parametrize("mode", ["graph", "blocked"])
def test_each_solve_returns_its_own_flow(T, small_case, mode):
    """solve_stereo returns a new tensor every call, as the reference returns a new array
    (Python/vfs.py:462-533): keeping the first result across a second solve of other
    images leaves it unchanged."""
    pair = small_case["pair"]
    v = _solver(small_case, mode)
    first = v.solve_stereo(pair["img_ref"], pair["img_other"])
    kept = first.clone()
    second = v.solve_stereo(pair["img_other"], pair["img_ref"])
    T.cuda.synchronize()
    assert first.data_ptr() != second.data_ptr()
    assert T.equal(first, kept)
    assert not T.equal(first, second)
    assert v.warpUV is second


def test_graph_replays_track_new_inputs(T, small_case):
    """Replays read the new images: eager solves interleaved with replays agree every time."""
    pair = small_case["pair"]
    g = _solver(small_case, "graph")
    e = _solver(small_case, "eager")
    rng = np.random.default_rng(0)
    for i in range(3):
        noise = rng.integers(-2, 3, pair["img_other"].shape)
        other = np.clip(pair["img_other"].astype(int) + noise, 0, 255).astype(np.uint8)
        fg = g.solve_stereo(pair["img_ref"], other).cpu().numpy()
        fe = e.solve_stereo(pair["img_ref"], other).cpu().numpy()
        T.cuda.synchronize()
        np.testing.assert_array_equal(fg, fe)


def test_depth_matches_oracle(T, small_case, small_oracle):
    import equidistant_cam as E
    pair = small_case["pair"]
    v = _solver(small_case, "graph")
    v.solve_stereo(pair["img_ref"], pair["img_other"])
    cam = E.EquidistantCam(pair["cam"])
    cam.setup_module()
    xyz = v.triangulate_flow(cam, pair["t"]).cpu().numpy()
    xyz_o = O.triangulate_flow(small_oracle["flow"], pair["cam"], pair["t"].reshape(3), pair["mask"])
    m = pair["mask"] > 0
    z, zo = xyz[..., 2][m], xyz_o[..., 2][m]
    ok = np.isfinite(zo) & (np.abs(zo) > 1e-3)
    rel = np.abs(z[ok] - zo[ok]) / np.abs(zo[ok])
    print(f"depth rel diff vs oracle: median {np.median(rel):.3e} p99 {np.percentile(rel, 99):.3e}")
    assert np.median(rel) <= 1e-3
    assert np.all(xyz[~m] == [0, 0, -1])


def test_residual_norm_matches_numpy_and_tracks_warps(T):
    """vfs_residual_norm (masked, multi-channel, deterministic) against NumPy, and the
    per-warp residuals of the Python-driven solve: one per warp and level, coarse to
    fine, finite and non-negative."""
    import torch
    import synthetic as S
    import vfs as V
    import vfs_params as P
    rng = np.random.default_rng(3)
    a = rng.normal(size=(37, 53, 2)).astype(np.float32)
    b = rng.normal(size=(37, 53, 2)).astype(np.float32)
    m = (rng.random((37, 53)) > 0.3).astype(np.uint8) * 255
    v = V.VFS(P.vfs_params(np.zeros((8, 8), np.uint8)))
    v.setup_module()
    ta, tb, tm = (torch.as_tensor(x, device="cuda") for x in (a, b, m))
    got = float(v.residual_norm(ta, tb, tm))
    ref = np.sqrt(np.sum(((a - b).astype(np.float64) ** 2)[m > 0]) / np.count_nonzero(m))
    assert abs(got - ref) <= 1e-6 * ref
    got1 = float(v.residual_norm(ta[..., 0].contiguous(), tb[..., 0].contiguous()))
    ref1 = np.sqrt(np.mean((a[..., 0] - b[..., 0]).astype(np.float64) ** 2))
    assert abs(got1 - ref1) <= 1e-6 * ref1
    assert float(v.residual_norm(ta, tb, tm)) == got  # deterministic
    # per-warp tracking in the Python-driven solve
    n = 64
    pr = S.make_pair(n, n, 2, t=(0.2, 0.0, 0.0), rot_deg=1.0, n_levels=2)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((n, n)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 4
    prm.Lambda, prm.alpha0, prm.alpha1, prm.beta, prm.gamma = 50.0, 5.0, 1.0, 10.0, 0.5
    v = V.VFS(prm)
    v.solver = "python"
    v.track_residuals = True
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pr["mask"])
    v.create_vector_field_pyramid(S.sanitize_field(tvf), cvf.astype(np.float32))
    v.solve_stereo(pr["img_ref"], pr["img_other"])
    assert [(lv, w) for lv, w, _ in v.residuals] == [(1, 0), (1, 1), (1, 2), (0, 0), (0, 1), (0, 2)]
    assert all(np.isfinite(r) and r >= 0 for _, _, r in v.residuals)
    assert v.residuals[0][2] > 0
