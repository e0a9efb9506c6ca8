"""Parameter fit (vfs-python_amd/fit.py): Adam on log(theta) with the adjoint's gradient."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_first_adam_step_moves_every_log_parameter_by_lr_against_the_gradient(T):
    import fit
    pairs = fit.scene_pairs(2, (64, 64), 11)
    f = fit.ParameterFit(pairs, fit.START, lr=0.01, reverse_warps=0, min_disparity=0.0)
    phi0 = f.phi.copy()
    loss, g = f.loss_and_grad([0, 1])
    assert np.isfinite(loss) and np.all(np.isfinite(g)) and np.all(g != 0)
    f.step([0, 1])
    # first Adam step: m_hat = g, v_hat = g^2, so the step is -lr g / (|g| + eps)
    np.testing.assert_allclose(f.phi - phi0, -0.01 * g / (np.abs(g) + f.eps), rtol=1e-9, atol=1e-15)
    h = f.run(2, 2)
    assert len(h) == 3 and all(np.isfinite(e["loss"]) for e in h)
    assert all(v > 0 for v in f.theta().values())


def test_median_depth_error_runs_on_held_out_pairs(T):
    """The evaluation solve (blocked path, the fit's schedule): one finite error per pair
    and their median.  At 96 x 96 the disparities are ~1 px, so the error itself is large."""
    import fit
    pairs = fit.scene_pairs(3, (96, 96), 5)
    err, per, loss = fit.median_depth_error(pairs, fit.START, min_disparity=0.0)
    assert len(per) == 3 and all(np.isfinite(e) and e >= 0 for e in per)
    assert err == float(np.median(per)) and np.isfinite(loss) and loss >= 0


def test_truncated_gradient_is_the_gradient_of_the_last_warp(T):
    """reverse_warps = n_warp reverses every warp of each level but holds the level's start
    state fixed: with one level this is the exact gradient."""
    import fit
    import vfs_adjoint as A
    import vfs_params as P
    pr = fit.scene_pairs(1, (64, 64), 4)[0]
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 1, 2.0, 3, 6
    for k, v in fit.START.items():
        setattr(prm, k, v)
    ref0 = torch.as_tensor(pr["img_ref"], device="cuda").to(torch.float32) / 256.0
    g = []
    for rw in (0, 3, 1):
        a = A.VFSDepthAdjoint(prm, pr["mask"], pr["tvf"], pr["cvf"], pr["cam"], pr["t"], pr["z_gt"], checkpoint=4,
                              reverse_warps=rw)
        a.forward(ref0, pr["img_other"])
        g.append(a.backward())
    for k in fit.PARAMS:
        assert abs(g[1][k] - g[0][k]) <= 1e-5 * abs(g[0][k]) + 1e-9, k
    assert any(abs(g[2][k] - g[0][k]) > 1e-6 * abs(g[0][k]) for k in fit.PARAMS)
