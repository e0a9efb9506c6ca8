"""Depth colouring (Python/colormap.py:25-292) and point-cloud export
(Python/main_equidistant.py:80-96), on the CPU.

The colour table is built from jet's definition; it is pinned by the reference's
own table, kept as data in tests/golden/colormap_table.npz (written by
tests/golden/make_colormap_fixture.py), and by jet's breakpoints.  depth_to_colormap
is pinned by hand-worked values at the ends of the range (index 255 at min_z:
0.5 * 255 -> 127 in red; index 0 at max_z: 0.5156 * 255 -> 131 in blue) and by
its monotonicity in inverse depth.  The PCD writer has no Open3D to compare with
(not installed): it is pinned by its byte layout and by round trips.
"""

import os

import numpy as np

import colormap as C
import point_cloud as PC

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "colormap_table.npz")


def test_table_matches_reference_table():
    ref = np.load(GOLDEN)["colormap_raw"]
    np.testing.assert_array_equal(C.colormap_raw, ref)
    np.testing.assert_array_equal(C.colormap, ref[:, ::-1])


def test_jet_breakpoints():
    j = C.jet(256)
    np.testing.assert_allclose(j[0], [0, 0, 33 / 64])
    np.testing.assert_allclose(j[31], [0, 0, 1])
    np.testing.assert_allclose(j[95], [0, 1, 1])
    np.testing.assert_allclose(j[159], [1, 1, 0])
    np.testing.assert_allclose(j[223], [1, 0, 0])
    np.testing.assert_allclose(j[255], [0.5, 0, 0])
    assert np.all(np.abs(np.diff(j, axis=0)) <= 1 / 64 + 1e-12)  # ramps of 64 steps


def test_depth_to_colormap_known_values():
    near, far = 1.5, 10.0
    d = np.array([[near, far, 0.0, -3.0, np.nan, 0.5, 1e9]])
    img = C.depth_to_colormap(d, near, far)
    assert img.dtype == np.uint8 and img.shape == (1, 7, 3)
    np.testing.assert_array_equal(img[0, 0], [0, 0, 127])    # nearest: dark red (B, G, R)
    np.testing.assert_array_equal(img[0, 1], [131, 0, 0])    # farthest: dark blue
    np.testing.assert_array_equal(img[0, 2:5], np.zeros((3, 3)))  # no depth: black
    np.testing.assert_array_equal(img[0, 5], img[0, 0])      # nearer than near clamps
    np.testing.assert_array_equal(img[0, 6], img[0, 1])      # farther than far clamps


def test_depth_to_colormap_monotone():
    near, far = 1.5, 10.0
    z = np.linspace(near, far, 2000)
    delta = 1 / near - 1 / far
    idx = np.round(np.clip(1 / (z + 1e-6) - 1 / far, 0, delta) / delta * 255).astype(int)
    assert np.all(np.diff(idx) <= 0) and idx[0] == 255 and idx[-1] == 0
    img = C.depth_to_colormap(z[None], near, far)[0]
    np.testing.assert_array_equal(img, (C.colormap[idx] * 255).astype(np.uint8))


def test_bgr_png(tmp_path):
    from PIL import Image
    img = np.random.default_rng(0).integers(0, 256, (5, 7, 3), dtype=np.uint8)
    C.write_bgr_png(tmp_path / "a.png", img)
    np.testing.assert_array_equal(np.array(Image.open(tmp_path / "a.png")), img[..., ::-1])


def test_pcd_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    xyz = rng.normal(size=(6, 8, 3)).astype(np.float32)
    img = rng.integers(0, 256, (6, 8, 3), dtype=np.uint8)
    pts, cols = PC.colored_points(xyz, img)
    np.testing.assert_array_equal(cols[3] * 255, img.reshape(-1, 3)[3, ::-1])  # R, G, B
    for binary in (True, False):
        p = tmp_path / f"c{int(binary)}.pcd"
        PC.write_pcd(p, pts, cols, binary=binary)
        head = p.read_bytes().split(b"DATA")[0].decode()
        assert "FIELDS x y z rgb" in head and "POINTS 48" in head and "TYPE F F F F" in head
        p2, c2 = PC.read_pcd(p)
        np.testing.assert_array_equal(p2, pts)
        np.testing.assert_array_equal(np.rint(c2 * 255), img.reshape(-1, 3)[:, ::-1])
    raw = (tmp_path / "c1.pcd").read_bytes()
    body = raw[raw.index(b"DATA binary\n") + len(b"DATA binary\n"):]
    assert len(body) == 48 * 16
    first = np.frombuffer(body[:16], dtype="<u4")
    b, g, r = (int(v) for v in img[0, 0])
    assert first[3] == (r << 16) | (g << 8) | b
    PC.write_pcd(tmp_path / "plain.pcd", pts)
    p3, c3 = PC.read_pcd(tmp_path / "plain.pcd")
    np.testing.assert_array_equal(p3, pts)
    assert c3 is None
