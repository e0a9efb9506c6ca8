"""The solve at the reference's default parameters (Python/vfs_params.py:25-43):
11 levels at fScale 1.2, lambda 5, alpha0 17, alpha1 1.2, beta 9, gamma 0.85, against
the float64 oracle, on a non-square image so every pyramid level has its own odd size.

Workload: seeded synthetic pair of 120 x 136 (H x W), f = 120/pi, t = (0.2, 0, 0) m,
1 deg rotation; the pyramid runs 120x136 ... 17x20.  The warp and iteration counts
are cut (the defaults are 100 x 100) so the float64 oracle solve stays within
seconds; every kernel and the 1.2 up/down-sampling factor still run on all 11
levels.

Tolerances.  At 2 warps x 6 iterations per level: median |d flow| <= 1e-4 px at
full resolution (the float32 restatement of the same formulas: median 3e-6 px,
p99 3.5e-4 px).  At 4 warps x 12 iterations the 44 warps down the pyramid amplify
rounding past that bar for any float32 evaluation (the float32 restatement:
median 6.4e-4 px, p99 0.045 px), so the GPU is held to <= 3x the restatement's
median and 99th percentile there.  Both: median relative depth difference <= 1e-3,
flow outside the mask exactly zero, and all solver modes agree bit for bit.
"""

import numpy as np
import pytest

from oracle import vfs_oracle as O
from conftest import flow_diff_stats

pytestmark = pytest.mark.gpu

H, W = 120, 136
STRICT = (2, 6)  # (warps, iterations) per level held to the absolute bar
LONG = (4, 12)    # held to the float32 restatement


@pytest.fixture(scope="module")
def case():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import synthetic as S
    pair = S.make_pair(H, W, 7, t=(0.2, 0.0, 0.0), rot_deg=1.0, n_levels=3)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((H, W)), pair["cam"], pair["cam"], pair["R"], pair["t"], 1)
    tvf = S.sanitize_field(tvf)
    cvf = np.asarray(cvf, dtype=np.float32)
    out = dict(pair=pair, tvf=tvf, cvf=cvf)
    for sched in (STRICT, LONG):
        prm = O.OracleParams(H, W, nWarpIters=sched[0], nSolverIters=sched[1])
        args = (pair["img_ref"], pair["img_other"], pair["mask"], tvf, cvf, prm)
        f64 = O.solve_stereo(*args)
        with O.precision(np.float32):
            f32 = O.solve_stereo(*args)
        out[sched] = dict(flow=f64, flow32=f32)
    return out


def _solve(case, mode, sched):
    import vfs as V
    import vfs_params as P
    pair = case["pair"]
    prm = P.vfs_params(pair["img_ref"])
    assert (prm.nLevel, prm.fScale, prm.Lambda, prm.alpha0, prm.alpha1, prm.beta, prm.gamma) == \
        (11, 1.2, 5.0, 17.0, 1.2, 9.0, 0.85)
    prm.nWarpIters, prm.nSolverIters = sched
    v = V.VFS(prm)
    v.solver = mode
    v.setup_module()
    v.compute_pyramid_sizes()
    assert len(v.sizes) == 11 and tuple(v.sizes[-1]) == (17, 20)
    v.create_mask_pyramid(pair["mask"])
    v.create_vector_field_pyramid(case["tvf"], case["cvf"])
    return v, v.solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()


@pytest.mark.parametrize("sched", [STRICT, LONG], ids=["2x6", "4x12"])
def test_default_params_match_oracle(case, sched):
    import equidistant_cam as E
    v, f = _solve(case, "graph", sched)
    mask = case["pair"]["mask"]
    ref = case[sched]
    st = flow_diff_stats(f, ref["flow"], mask)
    st32 = flow_diff_stats(ref["flow32"], ref["flow"], mask)
    print(f"default params {sched}, GPU vs oracle:", st)
    print(f"default params {sched}, fp32 restatement vs oracle:", st32)
    assert np.all(f[mask == 0] == 0)
    if sched == STRICT:
        assert st["median"] <= 1e-4
        assert st["p99"] <= max(3 * st32["p99"], 1e-4)
    else:
        assert st["median"] <= 3 * st32["median"]
        assert st["p99"] <= 3 * st32["p99"]
    pair = case["pair"]
    cam = E.EquidistantCam(pair["cam"])
    cam.setup_module()
    z = v.triangulate_flow(cam, pair["t"]).cpu().numpy()[..., 2]
    zo = O.triangulate_flow(ref["flow"], pair["cam"], pair["t"].reshape(3), mask)[..., 2]
    m = mask > 0
    ok = np.isfinite(zo[m]) & (np.abs(zo[m]) > 1e-3)
    rel = np.abs(z[m][ok] - zo[m][ok]) / np.abs(zo[m][ok])
    print(f"default params {sched} depth rel diff: median {np.median(rel):.3e}")
    assert np.median(rel) <= 1e-3


def test_default_params_modes_bitwise(case):
    outs = {mode: _solve(case, mode, LONG)[1] for mode in ("graph", "eager", "python", "blocked", "blocked_eager")}
    for mode in ("eager", "python", "blocked", "blocked_eager"):
        np.testing.assert_array_equal(outs[mode], outs["graph"], err_msg=mode)
