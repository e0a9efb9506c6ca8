"""CPU checks of the C ABI boundary and the host-side logic (no GPU compute calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vfs_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vfs_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    import vfs_hip as H
    assert _declared() == H.exported_symbols()


def test_library_loads_and_exports_every_declared_symbol():
    import vfs_hip as H
    if not os.path.exists(H.LIB_PATH):
        pytest.skip("libvfs_hip.so not built")
    L = ctypes.CDLL(H.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name
    assert H.lib().vfs_version().startswith(b"vfs_hip")


def test_invalid_arguments_are_rejected_before_any_launch():
    import vfs_hip as H
    if not os.path.exists(H.LIB_PATH):
        pytest.skip("libvfs_hip.so not built")
    L = H.lib()
    assert L.vfs_fill(None, None, 0.0, 10) == H.VFS_EINVAL
    assert L.vfs_update_dual(None, None, None, None, 1.0, None, None, None, None, 4, 4, None) == H.VFS_EINVAL
    assert L.vfs_median_filter(None, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(32), 4, 4, 5) == H.VFS_EINVAL
    assert L.vfs_upsample(None, ctypes.c_void_p(16), 0, 4, ctypes.c_void_p(32), 2, 2, 1, 2.0) == H.VFS_EINVAL
    assert L.vfs_solve_level_plain(None, None, None, None, None, None, 1, 1) == H.VFS_EINVAL
    with pytest.raises(RuntimeError):
        H.check(H.VFS_EINVAL, "x")


def test_params_struct_matches_reference_constants():
    import vfs_hip as H
    import vfs_params as P
    prm = P.vfs_params()
    k = H.make_params(prm)
    assert k.eta_p == 3.0 and k.eta_q == 2.0 and k.median_k == 5
    assert k.min_grad == np.float32(1e-8) and k.min_tensor == np.float32(1e-8)
    assert k.a11 == np.float32(1.2 * 1.2)  # computed in double, then rounded (Python/vfs.cu:231)
    assert k.a004 == np.float32(4 * 17.0 * 17.0)
    assert ctypes.sizeof(H.VfsParams) == 14 * 4
    assert ctypes.sizeof(H.VfsLevel) == 8 + 19 * 8


def test_vfs_params_defaults_match_reference():
    import vfs_params as P
    p = P.vfs_params()
    assert (p.width, p.height) == (848, 800)
    assert (p.beta, p.gamma, p.alpha0, p.alpha1, p.timeStepLambda, p.Lambda) == (9.0, 0.85, 17.0, 1.2, 1.0, 5.0)
    assert (p.nLevel, p.fScale, p.nWarpIters, p.nSolverIters, p.limitRange) == (11, 1.2, 100, 100, 0.1)
    img = np.zeros((30, 40))
    assert (P.vfs_params(img).height, P.vfs_params(img).width) == (30, 40)


def test_schedule_matches_oracle_and_closed_form():
    import vfs_hip as H
    from oracle import vfs_oracle as O
    t, s, m = H.pd_schedule(50, 1.0)
    to, so, mo = O.pd_schedule(50, 1.0)
    np.testing.assert_array_equal(t, to.astype(np.float32))
    np.testing.assert_array_equal(s, so.astype(np.float32))
    np.testing.assert_array_equal(m, mo.astype(np.float32))
    np.testing.assert_allclose(m, 1 / np.sqrt(1 + 0.7 * t.astype(np.float64)), rtol=1e-6)
    assert s[-1] == pytest.approx(18.4, rel=0.05)  # sigma after 50 iterations stays far below 1000


def test_pyramid_sizes_and_compute_optical_flow():
    import vfs as V
    import vfs_params as P
    from oracle import vfs_oracle as O
    prm = P.vfs_params()
    v = V.VFS(prm)
    v.compute_pyramid_sizes()
    assert v.sizes == O.compute_pyramid_sizes(800, 848, 11, 1.2)
    import torch
    u = torch.arange(6, dtype=torch.float32).reshape(2, 3)
    vec = torch.ones((2, 3, 2))
    assert torch.equal(V.VFS.compute_optical_flow(u, vec)[..., 1], u)


def test_texture_descriptor_semantics():
    import torch
    import texture
    t = torch.zeros((4, 5))
    to = texture.create_texture_object(t, addressMode=texture.cudaAddressModeMirror,
                                       filterMode=texture.cudaFilterModeLinear, normalizedCoords=1)
    t += 1  # the descriptor holds a snapshot, like the reference's cudaArray copy
    assert float(to.tensor.sum()) == 0.0
    assert to.shape == (4, 5)
    assert texture.create_texture_object(torch.zeros(3, dtype=torch.float64)) is None
    b = texture.create_texture_object(t)  # Border / Point / unnormalised defaults
    with pytest.raises(NotImplementedError):
        b.require(texture.cudaAddressModeMirror)


def test_synthetic_generator_is_seeded_and_masked():
    import synthetic as S
    a = S.make_pair(64, 64, 7, n_levels=3)
    b = S.make_pair(64, 64, 7, n_levels=3)
    np.testing.assert_array_equal(a["img_other"], b["img_other"])
    assert a["img_ref"].dtype == np.uint8 and a["mask"].dtype == np.uint8
    m = a["mask"]
    assert m[:5].max() == 0 and m[-5:].max() == 0 and m[:, :5].max() == 0 and m[:, -5:].max() == 0
    from oracle import vfs_oracle as O
    pyr = O.create_mask_pyramid(m, O.compute_pyramid_sizes(64, 64, 3, 2.0))
    for lvl in pyr:  # zero outer ring at every level
        assert lvl[0].max() == 0 and lvl[-1].max() == 0 and lvl[:, 0].max() == 0 and lvl[:, -1].max() == 0


def test_launch_count_and_graph_bound():
    """Launches per solve: the per-kernel solve issues 6 + 3 n_iter per warp, the
    blocked one 2 + ceil(n_iter / k); solves above the graph bound launch eagerly."""
    import vfs as V
    import vfs_params as P
    import synthetic as S
    prm = P.vfs_params(np.zeros((800, 800), np.uint8))
    v = V.VFS(prm)
    v.compute_pyramid_sizes()
    v.solver = "graph"
    n_plain = v.launch_count()
    assert n_plain == 4 * 11 + 4 + 11 * (3 + 100 * (6 + 3 * 100))
    assert n_plain > V.GRAPH_MAX_LAUNCHES
    v.solver = "blocked"
    v.pd_tiles = {lvl: (16, 16, 2) for lvl in range(11)}
    assert v.launch_count() == 4 * 11 + 4 + 11 * (3 + 100 * (2 + 50))
    v.pd_tiles = None
    assert v.launch_count() <= V.GRAPH_MAX_LAUNCHES  # the default parameters' blocked solve is recorded
    S.apply_workload_params(prm, "plain_full_800")
    v.compute_pyramid_sizes()
    v.solver = "graph"
    assert v.launch_count() <= V.GRAPH_MAX_LAUNCHES
