"""Per-kernel parity: HIP kernels (through the VFS / EquidistantCam API and the C ABI)
against the float64 oracle on the same float32 inputs.

Tolerances: integer/index work (mask pyramid, median selection) is bit-exact;
fp32 kernels are held to a few float32 ulps of the float64 oracle, stated per
test.
"""

import numpy as np
import pytest

from oracle import vfs_oracle as O

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (33, 65), (5, 7), (96, 128)]


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _vfs(Hh, Ww, **kw):
    import vfs as V
    import vfs_params as P
    prm = P.vfs_params()
    prm.height, prm.width = Hh, Ww
    prm.fScale = 2.0
    prm.nLevel = 2
    prm.Lambda, prm.alpha0, prm.alpha1, prm.beta, prm.gamma = 50.0, 5.0, 1.0, 10.0, 0.5
    for k, v in kw.items():
        setattr(prm, k, v)
    v = V.VFS(prm)
    v.setup_module()
    return v, prm


def _dev(T, a, dtype=None):
    a = np.ascontiguousarray(a)
    t = T.as_tensor(a, device="cuda")
    return t if dtype is None else t.to(dtype)


def _host(t):
    return t.detach().cpu().numpy()


def _mask(Hh, Ww, rng, p=0.15):
    m = (rng.random((Hh, Ww)) > p).astype(np.uint8) * 255
    return m


@pytest.mark.parametrize("hw", SIZES)
def test_cast_and_pyramids(T, hw):
    import texture
    rng = np.random.default_rng(1)
    Hh, Ww = hw[0] * 2, hw[1] * 2
    v, prm = _vfs(Hh, Ww)
    v.compute_pyramid_sizes()
    img = rng.integers(0, 256, (Hh, Ww), dtype=np.uint8)
    ref0 = O.cast_and_scale(img)
    v.calibration = T.zeros((Hh, Ww, 2), device="cuda")
    v.create_image_pyramid(img, img)
    np.testing.assert_array_equal(_host(v.refs[0]), ref0.astype(np.float32))
    np.testing.assert_allclose(_host(v.refs[1]), O.downsample_bilinear(ref0, v.sizes[1]), rtol=0, atol=1e-7)
    mask = _mask(Hh, Ww, rng)
    v.create_mask_pyramid(mask)
    np.testing.assert_array_equal(_host(v.masks[1]), O.downsample_mask(mask, v.sizes[1]))
    tvf = rng.normal(size=(Hh, Ww, 2)).astype(np.float32)
    v.create_vector_field_pyramid(tvf, np.zeros_like(tvf))
    np.testing.assert_allclose(_host(v.translations[1]), O.downsample_bilinear(tvf, v.sizes[1]), rtol=0, atol=1e-6)
    assert isinstance(v.masks_to[1], texture.SampledImage)


@pytest.mark.parametrize("hs,ws,hd,wd", [(800, 848, 666, 706), (25, 37, 12, 18), (11, 7, 9, 5)])
def test_mask_downsampling_general_ratio_bit_exact(T, hs, ws, hd, wd):
    import vfs_hip as H
    rng = np.random.default_rng(2)
    src = _mask(hs, ws, rng, 0.5)
    d = T.empty((hd, wd), dtype=T.uint8, device="cuda")
    s = _dev(T, src)
    H.call("vfs_downsample_mask", H.stream(), H.ptr(d), hd, wd, H.ptr(s), hs, ws)
    np.testing.assert_array_equal(_host(d), O.downsample_mask(src, (hd, wd)))


@pytest.mark.parametrize("hw", SIZES)
def test_warp_and_find_warping_vector(T, hw):
    import texture
    rng = np.random.default_rng(3)
    Hh, Ww = hw
    v, _ = _vfs(Hh, Ww)
    img = rng.random((Hh, Ww)).astype(np.float32)
    tvf = rng.normal(size=(Hh, Ww, 2)).astype(np.float32)
    flow = (rng.normal(size=(Hh, Ww, 2)) * 3).astype(np.float32)
    out = _host(v.warping_image(texture.wrap(_dev(T, img)), _dev(T, flow)))
    np.testing.assert_allclose(out, O.warping_image(img, flow), rtol=0, atol=2e-5)
    tv = _host(v.find_warping_vector(_dev(T, flow), texture.wrap(_dev(T, tvf))))
    np.testing.assert_allclose(tv, O.find_warping_vector(flow, tvf), rtol=0, atol=1e-4)


@pytest.mark.parametrize("hw", SIZES)
def test_gaussian_tensor_eta(T, hw):
    import texture
    rng = np.random.default_rng(4)
    Hh, Ww = hw
    v, prm = _vfs(Hh, Ww)
    img = rng.random((Hh, Ww)).astype(np.float32)
    mask = _mask(Hh, Ww, rng)
    tm = texture.wrap(_dev(T, mask), texture.cudaFilterModePoint)
    g = v.gaussian(texture.wrap(_dev(T, img)), tm, (Hh, Ww))
    g_o = O.gaussian_masked(img, mask)
    np.testing.assert_allclose(_host(g), g_o, rtol=0, atol=1e-7)
    abc, eta = v.compute_tensor_eta(g, tm)
    abc_o, eta_o = O.compute_tensor_eta(_host(g).astype(np.float64), mask, prm.beta, prm.gamma, prm.alpha0,
                                        prm.alpha1)
    np.testing.assert_allclose(_host(abc), abc_o, rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(_host(eta), eta_o, rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize("hw", SIZES)
def test_derivatives(T, hw):
    import texture
    rng = np.random.default_rng(5)
    Hh, Ww = hw
    v, _ = _vfs(Hh, Ww)
    ref = rng.random((Hh, Ww)).astype(np.float32)
    ow = rng.random((Hh, Ww)).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, (Hh, Ww))
    tvf = (np.stack([np.cos(ang), np.sin(ang)], -1) * rng.uniform(0.3, 2, (Hh, Ww, 1))).astype(np.float32)
    d = _host(v.compute_derivatives(_dev(T, ow), texture.wrap(_dev(T, ref)), texture.wrap(_dev(T, tvf))))
    d_o = O.compute_derivatives(ow, ref, tvf)
    np.testing.assert_allclose(d, d_o, rtol=0, atol=2e-5)


def _pd_inputs(rng, Hh, Ww):
    mask = _mask(Hh, Ww, rng)
    abc = rng.random((Hh, Ww, 3)).astype(np.float32)
    eta = (rng.random((Hh, Ww, 3)) * 20 + 1).astype(np.float32)
    return dict(mask=mask, abc=abc, eta=eta,
                p=(rng.normal(size=(Hh, Ww, 2)) * 0.3).astype(np.float32),
                q=(rng.normal(size=(Hh, Ww, 4)) * 0.3).astype(np.float32),
                u=(rng.normal(size=(Hh, Ww))).astype(np.float32),
                u_=(rng.normal(size=(Hh, Ww))).astype(np.float32),
                v_=(rng.normal(size=(Hh, Ww, 2)) * 0.1).astype(np.float32),
                deriv=(rng.normal(size=(Hh, Ww, 2)) * 0.2).astype(np.float32),
                tp=(rng.normal(size=(Hh, Ww, 2)) * 0.3).astype(np.float32))


@pytest.mark.parametrize("hw", SIZES)
@pytest.mark.parametrize("sigma", [1.0, 7.5])
def test_update_dual(T, hw, sigma):
    import texture
    rng = np.random.default_rng(6)
    Hh, Ww = hw
    v, prm = _vfs(Hh, Ww)
    a = _pd_inputs(rng, Hh, Ww)
    p, q = _dev(T, a["p"]), _dev(T, a["q"])
    tm = texture.wrap(_dev(T, a["mask"]), texture.cudaFilterModePoint)
    p, tp, q = v.update_dual_variables(p, q, sigma, _dev(T, a["u_"]), _dev(T, a["v_"]), _dev(T, a["abc"]), tm)
    po, tpo, qo = O.update_dual_variables(a["p"], a["q"], sigma, a["u_"], a["v_"], a["abc"], a["mask"],
                                          prm.alpha0, prm.alpha1)
    np.testing.assert_allclose(_host(p), po, rtol=0, atol=3e-6)
    np.testing.assert_allclose(_host(tp), tpo, rtol=0, atol=3e-6)
    np.testing.assert_allclose(_host(q), qo, rtol=0, atol=3e-6)


@pytest.mark.parametrize("hw", SIZES)
@pytest.mark.parametrize("tau", [1.0, 0.13])
def test_l1_thresholding(T, hw, tau):
    import texture
    rng = np.random.default_rng(7)
    Hh, Ww = hw
    v, prm = _vfs(Hh, Ww)
    a = _pd_inputs(rng, Hh, Ww)
    tm = texture.wrap(_dev(T, a["mask"]), texture.cudaFilterModePoint)
    out = v.l1_thresholding(tau, _dev(T, a["tp"]), _dev(T, a["u"]), _dev(T, a["u_"]), _dev(T, a["deriv"]),
                            _dev(T, a["eta"]), tm)
    ref = O.l1_thresholding(tau, a["tp"], a["u"], a["u_"], a["deriv"], a["eta"], a["mask"], prm.Lambda)
    np.testing.assert_allclose(_host(out), ref, rtol=1e-5, atol=2e-6)


@pytest.mark.parametrize("hw", SIZES)
def test_update_primal(T, hw):
    import texture
    rng = np.random.default_rng(8)
    Hh, Ww = hw
    v, prm = _vfs(Hh, Ww)
    a = _pd_inputs(rng, Hh, Ww)
    tm = texture.wrap(_dev(T, a["mask"]), texture.cudaFilterModePoint)
    tau, mu = 0.6, 0.8
    us, vs = v.update_primal_variables(tau, mu, _dev(T, a["p"]), _dev(T, a["q"]), _dev(T, a["u"]), _dev(T, a["u_"]),
                                       _dev(T, a["v_"]), _dev(T, a["abc"]), _dev(T, a["eta"]), tm)
    uo, vo = O.update_primal_variables(np.float32(tau), np.float32(mu), a["p"], a["q"], a["u"], a["u_"], a["v_"],
                                       a["abc"], a["eta"], a["mask"], prm.alpha0, prm.alpha1)
    np.testing.assert_allclose(_host(us), uo, rtol=0, atol=2e-6)
    np.testing.assert_allclose(_host(vs), vo, rtol=0, atol=2e-6)


@pytest.mark.parametrize("hw", SIZES)
@pytest.mark.parametrize("k", [5, 3])
def test_median_bit_exact(T, hw, k):
    import texture
    rng = np.random.default_rng(9)
    Hh, Ww = hw
    v, _ = _vfs(Hh, Ww)
    v.median_filter_kernel_size = k
    u = rng.normal(size=(Hh, Ww)).astype(np.float32)
    mask = _mask(Hh, Ww, rng)
    out = v.median_filter(_dev(T, u), texture.wrap(_dev(T, mask), texture.cudaFilterModePoint))
    np.testing.assert_array_equal(_host(out), O.median_filter(u, mask, k).astype(np.float32))


@pytest.mark.parametrize("hw", SIZES)
def test_flow_clamp_in_place(T, hw):
    import texture
    rng = np.random.default_rng(10)
    Hh, Ww = hw
    v, prm = _vfs(Hh, Ww)
    mask = _mask(Hh, Ww, rng)
    flow = rng.normal(size=(Hh, Ww, 2)).astype(np.float32)
    u0 = rng.normal(size=(Hh, Ww)).astype(np.float32)
    uf = (u0 + rng.normal(size=(Hh, Ww)) * 0.3).astype(np.float32)
    tv = rng.normal(size=(Hh, Ww, 2)).astype(np.float32)
    fd, ud = _dev(T, flow), _dev(T, u0)
    f2, u2 = v.compute_optical_flow_with_clamp(fd, ud, _dev(T, uf), _dev(T, tv),
                                               texture.wrap(_dev(T, mask), texture.cudaFilterModePoint))
    assert f2.data_ptr() == fd.data_ptr() and u2.data_ptr() == ud.data_ptr()
    fo, uo = O.compute_optical_flow_with_clamp(flow, u0, uf, tv, mask, prm.limitRange)
    np.testing.assert_allclose(_host(u2), uo, rtol=0, atol=1e-6)
    np.testing.assert_allclose(_host(f2), fo, rtol=0, atol=2e-6)


@pytest.mark.parametrize("hw", [(32, 32), (25, 50), (3, 4)])
def test_upsampling(T, hw):
    rng = np.random.default_rng(11)
    Hh, Ww = hw
    v, prm = _vfs(Hh * 2, Ww * 2)
    u = rng.normal(size=(Hh, Ww)).astype(np.float32)
    w = rng.normal(size=(Hh, Ww, 2)).astype(np.float32)
    out1 = _host(v.upsampling_vector_field1(_dev(T, u), (2 * Hh, 2 * Ww)))
    out2 = _host(v.upsampling_vector_field(_dev(T, w), (2 * Hh, 2 * Ww)))
    np.testing.assert_allclose(out1, O.upsampling(u, (2 * Hh, 2 * Ww), 2.0), rtol=0, atol=2e-6)
    np.testing.assert_allclose(out2, O.upsampling(w, (2 * Hh, 2 * Ww), 2.0), rtol=0, atol=2e-6)


def test_camera_kernels_and_vector_fields(T):
    import equidistant_cam as E
    import vector_field as VF
    Hh, Ww = 96, 128
    cam = (Ww / np.pi, Ww / 2.0, Hh / 2.0)
    c = E.EquidistantCam(cam)
    c.setup_module()
    rays = _host(c.get_rays((Hh, Ww)))
    np.testing.assert_allclose(rays, O.get_rays((Hh, Ww), *cam), rtol=0, atol=2e-6)
    rng = np.random.default_rng(12)
    xyz = rng.normal(size=(Hh, Ww, 3)).astype(np.float32)
    uv = _host(c.get_UVs(_dev(T, xyz)))
    np.testing.assert_allclose(uv, O.project(xyz, *cam), rtol=0, atol=2e-4)
    g = (rng.random((Hh, Ww, 2)) * [Ww, Hh]).astype(np.float32)
    np.testing.assert_allclose(_host(c.get_XYZs(_dev(T, g))), O.unproject(g, *cam), rtol=0, atol=2e-6)
    th = _host(c.get_thetas())
    r = O.get_rays((Hh, Ww), *cam)
    np.testing.assert_allclose(th[..., 2], np.arctan2(np.hypot(r[..., 0], r[..., 1]), r[..., 2]), atol=2e-6)
    # vector fields (the reference's only test regenerates these fields, Python/test_vector_field.py:52-66)
    ang = np.radians(1.5)
    R = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]])
    t = np.array([[0.3], [0.01], [0.02]])
    tvf, cvf = VF.get_vector_fields(VF.get_grid_zero((Hh, Ww)), c, c, R, t, 1, VF.unit_normalizer)
    tvf_o, cvf_o = O.get_vector_fields(O.get_grid_zero((Hh, Ww)), cam, cam, R, t, 1)
    ok = np.isfinite(tvf_o).all(-1) & (np.hypot(*(O.get_grid_zero((Hh, Ww)) - [Ww / 2, Hh / 2]).T).T < Hh / 2 - 2)
    assert np.max(np.abs(tvf - tvf_o)[ok]) < 5e-4
    assert np.max(np.abs(cvf - cvf_o)[ok]) < 5e-4
    tvf1, _ = VF.get_vector_fields(VF.get_grid_one((Hh, Ww)), c, c, R, t, (1, 0.01), VF.unit_normalizer)
    tvf1_o, _ = O.get_vector_fields(O.get_grid_zero((Hh, Ww)) + 1, cam, cam, R, t, (1, 0.01))
    assert np.max(np.abs(tvf1 - tvf1_o)[ok]) < 5e-4


def test_triangulation_and_distance(T):
    import equidistant_cam as E
    import vfs_hip as H
    Hh, Ww = 64, 64
    cam = (Ww / np.pi, Ww / 2.0, Hh / 2.0)
    rng = np.random.default_rng(13)
    g = O.get_grid_zero((Hh, Ww))
    D = 2 + 3 * rng.random((Hh, Ww))
    t = np.array([0.25, 0.0, 0.03])
    P = D[..., None] * O.unproject(g, *cam)
    flow = (O.project(P + t, *cam) - g).astype(np.float32)
    mask = ((np.hypot(g[..., 0] - Ww / 2, g[..., 1] - Hh / 2) < 25).astype(np.uint8) * 255)
    v, _ = _vfs(Hh, Ww)
    v.masks = [_dev(T, mask)]
    import texture
    v.masks_to = [texture.wrap(v.masks[0], texture.cudaFilterModePoint)]
    v.warpUV = _dev(T, flow)
    c = E.EquidistantCam(cam)
    c.setup_module()
    xyz = _host(v.triangulate_flow(c, t.reshape(3, 1)))
    ref = O.triangulate_flow(flow.astype(np.float64), cam, t, mask)
    m = mask > 0
    np.testing.assert_allclose(xyz[m], ref[m], rtol=2e-4, atol=1e-4)
    assert np.all(xyz[~m] == [0, 0, -1])
    c.set_mask(mask)
    xd = _host(c.get_xyz_from_distance(_dev(T, D.astype(np.float32))))
    np.testing.assert_allclose(xd[m], P[m], rtol=2e-6, atol=2e-6)


def test_derivatives_with_zero_length_translation_vectors(T):
    """A zero translation vector has no direction (hypot 0, Python/vfs.cu:127-129): Iu is
    non-finite there on both sides, Iz (a plain difference) stays finite, every other
    pixel matches, and no sample index leaves the image."""
    import texture
    rng = np.random.default_rng(21)
    Hh, Ww = 33, 41
    v, _ = _vfs(Hh, Ww)
    ref = rng.random((Hh, Ww)).astype(np.float32)
    ow = rng.random((Hh, Ww)).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, (Hh, Ww))
    tvf = np.stack([np.cos(ang), np.sin(ang)], -1).astype(np.float32)
    zero = rng.random((Hh, Ww)) < 0.05
    zero[0, 0] = zero[Hh - 1, Ww - 1] = zero[Hh // 2, Ww // 2] = True
    tvf[zero] = 0.0
    d = _host(v.compute_derivatives(_dev(T, ow), texture.wrap(_dev(T, ref)), texture.wrap(_dev(T, tvf))))
    d_o = O.compute_derivatives(ow, ref, tvf)
    assert np.array_equal(np.isfinite(d), np.isfinite(d_o))
    assert not np.isfinite(d[..., 0][zero]).any() and np.isfinite(d[..., 1]).all()
    ok = np.isfinite(d_o)
    np.testing.assert_allclose(d[ok], d_o[ok], rtol=0, atol=2e-5)


def test_triangulation_with_zero_disparity(T):
    """Where the other view's ray equals the reference ray (zero disparity, Python/vfs.cu:733)
    the depth has a zero denominator: the GPU gives the same non-finite values (sign
    included) as the oracle there, and matches it elsewhere."""
    import equidistant_cam as E
    import texture
    Hh, Ww = 48, 48
    cam = (Ww / np.pi, Ww / 2.0, Hh / 2.0)
    rng = np.random.default_rng(22)
    g = O.get_grid_zero((Hh, Ww))
    D = 2 + 3 * rng.random((Hh, Ww))
    t = np.array([0.25, 0.0, 0.03])
    P = D[..., None] * O.unproject(g, *cam)
    flow = (O.project(P + t, *cam) - g).astype(np.float32)
    still = rng.random((Hh, Ww)) < 0.1
    flow[still] = 0.0
    mask = ((np.hypot(g[..., 0] - Ww / 2, g[..., 1] - Hh / 2) < 20).astype(np.uint8) * 255)
    v, _ = _vfs(Hh, Ww)
    v.masks = [_dev(T, mask)]
    v.masks_to = [texture.wrap(v.masks[0], texture.cudaFilterModePoint)]
    v.warpUV = _dev(T, flow)
    c = E.EquidistantCam(cam)
    c.setup_module()
    xyz = _host(v.triangulate_flow(c, t.reshape(3, 1))).astype(np.float64)
    ref = O.triangulate_flow(flow.astype(np.float64), cam, t, mask)
    m = mask > 0
    assert (still & m).sum() > 10
    z, zo = xyz[..., 2][m], ref[..., 2][m]
    assert np.array_equal(np.isnan(z), np.isnan(zo))
    assert np.array_equal(np.isposinf(z), np.isposinf(zo)) and np.array_equal(np.isneginf(z), np.isneginf(zo))
    assert not np.isfinite(z[still[m]]).all()
    ok = np.isfinite(zo)
    np.testing.assert_allclose(z[ok], zo[ok], rtol=2e-4, atol=1e-4)
