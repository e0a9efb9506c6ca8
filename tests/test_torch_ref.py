"""The autograd-able PyTorch restatement (oracle/vfs_torch_ref.py) against the NumPy oracle.

The torch restatement is the checker of the hand-written adjoint; this pins its
forward to the float64 oracle (itself pinned in test_oracle.py).
"""

import numpy as np
import pytest
import torch

from oracle import vfs_oracle as O
from oracle import vfs_torch_ref as R


def _case(n=48, seed=11):
    import synthetic as S
    pr = S.make_pair(n, n, seed, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=2)
    cam = pr["cam"]
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((n, n)), cam, cam, pr["R"], pr["t"], 1)
    return pr, S.sanitize_field(tvf), cvf.astype(np.float32)


def test_torch_forward_matches_oracle():
    pr, tvf, cvf = _case()
    n = pr["img_ref"].shape[0]
    prm = O.OracleParams(n, n, nLevel=2, fScale=2.0, nWarpIters=2, nSolverIters=6, Lambda=50.0, alpha0=5.0,
                         alpha1=1.0, beta=10.0, gamma=0.5)
    fo = O.solve_stereo(pr["img_ref"], pr["img_other"], pr["mask"], tvf, cvf, prm)
    sizes = O.compute_pyramid_sizes(n, n, 2, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]))
    other_w = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf))
    t = lambda v: torch.tensor(v, dtype=torch.float64)  # noqa: E731
    ft = R.solve_stereo(ref0, other_w, masks, trs, sizes, t(50.0), t(5.0), t(1.0), t(10.0), t(0.5), 2, 6)
    np.testing.assert_allclose(ft.numpy(), fo, rtol=0, atol=1e-10)
    # triangulated depth agrees with the oracle's triangulation
    z = R.triangulate_z(ft, pr["cam"], torch.as_tensor(pr["t"].reshape(3)), masks[0]).numpy()
    zo = O.triangulate_flow(fo, pr["cam"], pr["t"].reshape(3), pr["mask"])[..., 2]
    m = pr["mask"] > 0
    ok = m & np.isfinite(zo) & (np.abs(zo) < 1e6)
    np.testing.assert_allclose(z[ok], zo[ok], rtol=1e-7)


def test_torch_gradients_are_finite():
    pr, tvf, cvf = _case(32, 3)
    n = 32
    sizes = O.compute_pyramid_sizes(n, n, 2, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"])).requires_grad_(True)
    other_w = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf))
    th = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (50.0, 5.0, 1.0, 10.0, 0.5)]
    flow = R.solve_stereo(ref0, other_w, masks, trs, sizes, *th, 2, 5)
    loss = R.depth_loss(flow, pr["cam"], torch.as_tensor(pr["t"].reshape(3)), masks[0],
                        torch.as_tensor(pr["z_gt"]))
    loss.backward()
    for g in [ref0.grad] + [x.grad for x in th]:
        assert g is not None and torch.isfinite(g).all()


def _oracle_depth_loss(pr, tvf, cvf, theta, n_warp, n_iter):
    n = pr["img_ref"].shape[0]
    prm = O.OracleParams(n, n, nLevel=2, fScale=2.0, nWarpIters=n_warp, nSolverIters=n_iter, **theta)
    flow = O.solve_stereo(pr["img_ref"], pr["img_other"], pr["mask"], tvf, cvf, prm)
    z = O.triangulate_flow(flow, pr["cam"], pr["t"].reshape(3), pr["mask"])[..., 2]
    m = (pr["mask"] > 0) & (z > 0) & np.isfinite(z) & (pr["z_gt"] > 0)
    return float(np.mean(np.abs(np.log(z[m]) - np.log(pr["z_gt"][m]))))


def test_torch_gradients_match_oracle_central_differences():
    """Autograd of the restatement is the adjoint's checker: pin it to central
    differences (h = 1e-4 relative) of the float64 NumPy oracle's depth loss."""
    pr, tvf, cvf = _case(40, 9)
    n, nw, ni = 40, 2, 5
    theta = dict(Lambda=50.0, alpha0=5.0, alpha1=1.0, beta=10.0, gamma=0.5)
    sizes = O.compute_pyramid_sizes(n, n, 2, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]))
    other_w = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf))
    names = tuple(theta)
    th = [torch.tensor(theta[k], dtype=torch.float64, requires_grad=True) for k in names]
    flow = R.solve_stereo(ref0, other_w, masks, trs, sizes, *th, nw, ni)
    loss = R.depth_loss(flow, pr["cam"], torch.as_tensor(pr["t"].reshape(3)), masks[0], torch.as_tensor(pr["z_gt"]))
    loss.backward()
    assert abs(float(loss) - _oracle_depth_loss(pr, tvf, cvf, theta, nw, ni)) <= 1e-10
    for i, k in enumerate(names):
        h = 1e-4 * theta[k]
        tp, tm = dict(theta), dict(theta)
        tp[k] += h
        tm[k] -= h
        fd = (_oracle_depth_loss(pr, tvf, cvf, tp, nw, ni) - _oracle_depth_loss(pr, tvf, cvf, tm, nw, ni)) / (2 * h)
        assert abs(th[i].grad.item() - fd) <= 5e-3 * abs(fd) + 1e-7, (k, th[i].grad.item(), fd)
