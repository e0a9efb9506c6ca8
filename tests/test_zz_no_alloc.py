"""The allocation-free entry point vfs_residual_norm_noalloc against vfs_residual_norm,
through the C ABI: the same bits on two calls in a row (the second with a smaller
input, so the static scratch holds the first call's partials when it starts), the
maximum number of partials accepted and one more rejected."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pair(T, shape, seed):
    rng = np.random.default_rng(seed)
    a = T.as_tensor(rng.normal(size=shape).astype(np.float32), device="cuda")
    b = T.as_tensor(rng.normal(size=shape).astype(np.float32), device="cuda")
    m = T.as_tensor((rng.random(shape[:2]) > 0.3).astype(np.uint8) * 255, device="cuda")
    return a, b, m


def _both(T, a, b, m, nparts):
    import vfs_hip as H
    L = H.lib()
    n = a.shape[0] * a.shape[1]
    ch = 1 if a.dim() == 2 else a.shape[2]
    s = H.stream()
    ref = T.empty(2, dtype=T.float64, device="cuda")
    part = T.empty(2 * nparts, dtype=T.float64, device="cuda")
    got = T.full((2,), -1.0, dtype=T.float64, device="cuda")
    mp = H.ptr(m) if m is not None else None
    assert L.vfs_residual_norm(s, H.ptr(ref), H.ptr(part), nparts, H.ptr(a), H.ptr(b), mp, n, ch) == 0
    assert L.vfs_residual_norm_noalloc(s, H.ptr(got), nparts, H.ptr(a), H.ptr(b), mp, n, ch) == 0
    T.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_residual_matches_workspace_entry_point(T):
    import vfs_hip as H
    cases = [((300, 411, 2), 1024, True), ((37, 53, 2), 8, True), ((64, 96), 24, False)]
    for shape, nparts, masked in cases:
        a, b, m = _pair(T, shape, sum(shape))
        ref, got = _both(T, a, b, m if masked else None, nparts)
        print(shape, nparts, ref, got, flush=True)
        assert got.tobytes() == ref.tobytes()
        # against NumPy (double sums of float32 squares)
        an, bn = a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)
        d2 = (an - bn) ** 2
        d2 = d2.sum(axis=2) if d2.ndim == 3 else d2
        sel = (m.cpu().numpy() > 0) if masked else np.ones(d2.shape, bool)
        assert got[1] == sel.sum()
        assert abs(got[0] - d2[sel].sum()) <= 1e-6 * d2[sel].sum()
    assert H.VFS_NOALLOC_MAX_PARTS == 1024


def test_noalloc_rejects_more_partials_than_its_scratch(T):
    import vfs_hip as H
    a, b, m = _pair(T, (16, 16), 1)
    out = T.zeros(2, dtype=T.float64, device="cuda")
    L = H.lib()
    rc = L.vfs_residual_norm_noalloc(H.stream(), H.ptr(out), H.VFS_NOALLOC_MAX_PARTS + 1, H.ptr(a), H.ptr(b),
                                     H.ptr(m), 256, 1)
    assert rc == H.VFS_ETOOBIG
    with pytest.raises(RuntimeError, match="capacity"):
        H.call("vfs_residual_norm_noalloc", H.stream(), H.ptr(out), H.VFS_NOALLOC_MAX_PARTS + 1, H.ptr(a),
               H.ptr(b), H.ptr(m), 256, 1)
    assert L.vfs_residual_norm_noalloc(H.stream(), H.ptr(out), 0, H.ptr(a), H.ptr(b), H.ptr(m), 256, 1) == H.VFS_EINVAL
    T.cuda.synchronize()
    assert out.cpu().tolist() == [0.0, 0.0]
