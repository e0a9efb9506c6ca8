"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces the committed outputs (a regression pin of the
oracle).  GPU: the HIP kernels and the full solve match the committed outputs.
"""

import os

import numpy as np
import pytest

from oracle import vfs_oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(G, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_kernel_fixture():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(G, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fx = _load("kernels_24x32.npz")
    a = mg.kernel_inputs()
    for k, v in a.items():
        np.testing.assert_array_equal(fx[k], v, err_msg=k)
    out = mg.kernel_outputs(a)
    for k, v in out.items():
        np.testing.assert_allclose(fx[k], v, rtol=1e-13, atol=1e-13, err_msg=k)


def test_oracle_reproduces_solve_fixture():
    fx = _load("solve_64.npz")
    n = fx["img_ref"].shape[0]
    lv, nw, ni = (int(x) for x in fx["schedule"])
    prm = O.OracleParams(n, n, nLevel=lv, fScale=2.0, nWarpIters=nw, nSolverIters=ni, alpha0=5.0, alpha1=1.0,
                         beta=10.0, gamma=0.5, Lambda=50.0, limitRange=0.1)
    flow = O.solve_stereo(fx["img_ref"], fx["img_other"], fx["mask"], fx["tvf"], fx["cvf"], prm)
    np.testing.assert_allclose(flow, fx["flow"], rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_hip_solve_matches_solve_fixture():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import vfs as V
    import vfs_params as P
    fx = _load("solve_64.npz")
    lv, nw, ni = (int(x) for x in fx["schedule"])
    prm = P.vfs_params(fx["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = lv, 2.0, nw, ni
    prm.alpha0, prm.alpha1, prm.beta, prm.gamma, prm.Lambda = 5.0, 1.0, 10.0, 0.5, 50.0
    v = V.VFS(prm)
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(fx["mask"])
    v.create_vector_field_pyramid(fx["tvf"], fx["cvf"])
    flow = v.solve_stereo(fx["img_ref"], fx["img_other"]).cpu().numpy()
    m = fx["mask"] > 0
    d = np.linalg.norm(flow.astype(np.float64) - fx["flow"], axis=-1)[m]
    assert np.median(d) <= 1e-4 and np.percentile(d, 99) <= 1e-3
    assert np.all(flow[~m] == 0)


@pytest.mark.gpu
def test_hip_kernels_match_kernel_fixture():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import vfs_hip as H
    import ctypes
    fx = _load("kernels_24x32.npz")
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    host = lambda t: t.cpu().numpy()  # noqa: E731

    class Prm:
        beta, gamma, alpha0, alpha1, Lambda, limitRange, fScale = 10.0, 0.5, 5.0, 1.0, 50.0, 0.1, 2.0

    k = H.make_params(Prm)
    s = H.stream()
    Hh, Ww = fx["mask"].shape
    mask = dev(fx["mask"])
    g = torch.empty((Hh, Ww), device="cuda")
    img = dev(fx["img"])
    H.call("vfs_gaussian_masked", s, H.ptr(g), H.ptr(img), H.ptr(mask), Hh, Ww)
    np.testing.assert_allclose(host(g), fx["gaussian"], atol=1e-7)
    abc = torch.empty((Hh, Ww, 3), device="cuda")
    eta = torch.empty((Hh, Ww, 3), device="cuda")
    H.call("vfs_compute_tensor_eta", s, H.ptr(abc), H.ptr(eta), H.ptr(g), H.ptr(mask), Hh, Ww, ctypes.byref(k))
    np.testing.assert_allclose(host(abc), fx["abc"], rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(host(eta), fx["eta"], rtol=2e-5, atol=2e-5)
    # keep every input tensor referenced until the call returns (a temporary passed
    # through H.ptr() could be freed and its block reused before the launch)
    p, q = dev(fx["p"]), dev(fx["q"])
    u_, v_, abc32 = dev(fx["u_"]), dev(fx["v_"]), dev(fx["abc"].astype(np.float32))
    tp = torch.empty_like(p)
    H.call("vfs_update_dual", s, H.ptr(p), H.ptr(tp), H.ptr(q), float(fx["sigma"]), H.ptr(u_), H.ptr(v_),
           H.ptr(abc32), H.ptr(mask), Hh, Ww, ctypes.byref(k))
    np.testing.assert_allclose(host(p), fx["p_out"], atol=3e-6)
    np.testing.assert_allclose(host(tp), fx["tp_out"], atol=3e-6)
    np.testing.assert_allclose(host(q), fx["q_out"], atol=3e-6)
    out = torch.empty((Hh, Ww), device="cuda")
    tpi, u, deriv, eta_in = dev(fx["tp"]), dev(fx["u"]), dev(fx["deriv"]), dev(fx["eta_in"])
    H.call("vfs_l1_thresholding", s, H.ptr(out), float(fx["tau"]), H.ptr(tpi), H.ptr(u), H.ptr(u_), H.ptr(deriv),
           H.ptr(eta_in), H.ptr(mask), Hh, Ww, ctypes.byref(k))
    np.testing.assert_allclose(host(out), fx["u_th"], rtol=1e-5, atol=2e-6)
    pi, qi, vb = dev(fx["p"]), dev(fx["q"]), dev(fx["v_"])
    ub_out, vb_out = torch.empty_like(u), torch.empty_like(vb)
    H.call("vfs_update_primal", s, H.ptr(ub_out), H.ptr(vb_out), float(fx["tau"]), float(fx["mu"]), H.ptr(pi),
           H.ptr(qi), H.ptr(u), H.ptr(u_), H.ptr(vb), H.ptr(abc32), H.ptr(eta_in), H.ptr(mask), Hh, Ww,
           ctypes.byref(k))
    np.testing.assert_allclose(host(ub_out), fx["ubar_out"], atol=2e-6)
    np.testing.assert_allclose(host(vb_out), fx["vbar_out"], atol=2e-6)
    md = torch.empty((Hh, Ww), device="cuda")
    H.call("vfs_median_filter", s, H.ptr(md), H.ptr(u), H.ptr(mask), Hh, Ww, 5)
    np.testing.assert_array_equal(host(md), fx["median"].astype(np.float32))
    mdn = torch.empty((12, 16), dtype=torch.uint8, device="cuda")
    H.call("vfs_downsample_mask", s, H.ptr(mdn), 12, 16, H.ptr(mask), Hh, Ww)
    np.testing.assert_array_equal(host(mdn), fx["mask_down"])
