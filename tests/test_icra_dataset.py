"""The ICRA dataset reader (Python/vfs_icra_dataset.py:30-89) and the image
preparation of Python/main_equidistant.py:50-67, on the CPU.

The dataset itself is not part of the reference, so a synthetic dataset of the
same layout is written to a temporary folder: poses in OpenCV FileStorage XML,
8-bit colour views, 16-bit range PNGs (metres x 256) and Middlebury .flo vector
fields.  The BGR-to-gray and histogram-equalisation restatements are pinned by
known answers of OpenCV's published fixed-point luma (pure blue 29, green 150,
red 76) and by the equalisation's defining properties; the disc mask by its
integer area against a brute-force count.  The known-answer test against the
dataset's own vector fields is in test_icra_kat_gpu.py.
"""

import os

import numpy as np
import pytest

import vfs_icra_dataset as D

XML = """<?xml version="1.0"?>
<opencv_storage>
<R type_id="opencv-matrix">
  <rows>3</rows>
  <cols>3</cols>
  <dt>d</dt>
  <data>
    {R}</data></R>
<t type_id="opencv-matrix">
  <rows>3</rows>
  <cols>1</cols>
  <dt>d</dt>
  <data>
    {t}</data></t>
</opencv_storage>
"""


def _fmt(a):
    return "\n    ".join(" ".join(f"{v:.17e}" for v in row) for row in np.atleast_2d(a).reshape(-1, 3))


@pytest.fixture()
def dataset(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(11)
    H, W = 24, 32
    frames = []
    for i, stem in enumerate(["0000000001", "0000000000"]):  # written out of order: the reader sorts
        c, s = np.cos(0.1 * (i + 1)), np.sin(0.1 * (i + 1))
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        t = rng.normal(size=(3, 1))
        rgb_ref = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        rgb_other = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        depth16 = rng.integers(0, 65536, (H, W)).astype(np.uint16)
        result16 = rng.integers(0, 65536, (H, W)).astype(np.uint16)
        tvf = rng.normal(size=(H, W, 2)).astype(np.float32)
        cvf = rng.normal(size=(H, W, 2)).astype(np.float32)
        for sub in ("calib", "image_02/data", "image_03/data", "proj_depth/groundtruth/image_02", "result_ours",
                    "translationVector", "calibrationVector"):
            os.makedirs(tmp_path / sub, exist_ok=True)
        (tmp_path / "calib" / f"{stem}.xml").write_text(
            XML.format(R=_fmt(R), t=" ".join(f"{v:.17e}" for v in t.ravel())))
        Image.fromarray(rgb_ref).save(tmp_path / "image_02/data" / f"{stem}.png")
        Image.fromarray(rgb_other).save(tmp_path / "image_03/data" / f"{stem}.png")
        Image.fromarray(depth16).save(tmp_path / "proj_depth/groundtruth/image_02" / f"{stem}.png")
        Image.fromarray(result16).save(tmp_path / "result_ours" / f"{stem}depth.png")
        D.write_flo(tmp_path / "translationVector" / f"{stem}.flo", tvf)
        D.write_flo(tmp_path / "calibrationVector" / f"{stem}.flo", cvf)
        frames.append(dict(stem=stem, R=R, t=t, rgb_ref=rgb_ref, rgb_other=rgb_other, depth16=depth16,
                           result16=result16, tvf=tvf, cvf=cvf))
    frames.sort(key=lambda f: f["stem"])
    return str(tmp_path), frames


def test_loader_reads_every_part(dataset):
    base, frames = dataset
    ld = D.dataset_loader(base)
    assert [os.path.basename(p) for p in ld.fn_xmls] == ["0000000000.xml", "0000000001.xml"]
    for i, fr in enumerate(frames):
        R, t = ld.load_pose(i)
        assert R.dtype == np.float64 and R.shape == (3, 3) and t.shape == (3, 1)
        np.testing.assert_array_equal(R, fr["R"])
        np.testing.assert_array_equal(t, fr["t"])
        ref, other = ld.load_image(i)
        np.testing.assert_array_equal(ref, fr["rgb_ref"][..., ::-1])  # B, G, R order
        np.testing.assert_array_equal(other, fr["rgb_other"][..., ::-1])
        d = ld.load_depth(i)
        assert d.dtype == np.float32
        np.testing.assert_array_equal(d, (fr["depth16"] / 256).astype(np.float32))
        np.testing.assert_array_equal(ld.load_result(i), (fr["result16"] / 256).astype(np.float32))
        tvf, cvf = ld.load_vector_field(i)
        np.testing.assert_array_equal(tvf, fr["tvf"])
        np.testing.assert_array_equal(cvf, fr["cvf"])


def test_calibration():
    f, cx, cy = D.dataset_loader("/nonexistent").get_calibration()
    assert (cx, cy) == (400, 400)
    assert abs(f * (2 * np.pi / 3) - 800.0) < 1e-9  # a ray at 60 deg lands 400 px from the centre
    assert abs(f * np.pi / 3 - 400.0) < 1e-9


def test_flo_format(tmp_path):
    flow = np.arange(2 * 3 * 5, dtype=np.float32).reshape(3, 5, 2)
    p = tmp_path / "a.flo"
    D.write_flo(p, flow)
    raw = p.read_bytes()
    assert raw[:4] == b"PIEH" and int.from_bytes(raw[4:8], "little") == 5 and int.from_bytes(raw[8:12], "little") == 3
    assert len(raw) == 12 + 4 * flow.size
    np.testing.assert_array_equal(D.read_flo(p), flow)
    (tmp_path / "bad.flo").write_bytes(b"XXXX" + raw[4:])
    with pytest.raises(ValueError):
        D.read_flo(tmp_path / "bad.flo")
    (tmp_path / "short.flo").write_bytes(raw[:-4])
    with pytest.raises(ValueError):
        D.read_flo(tmp_path / "short.flo")


def test_opencv_matrix_types(tmp_path):
    p = tmp_path / "m.xml"
    p.write_text('<?xml version="1.0"?>\n<opencv_storage>\n<K type_id="opencv-matrix"><rows>2</rows><cols>2</cols>'
                 '<dt>f</dt><data>1. 2.5 -3 4e-1</data></K>\n</opencv_storage>\n')
    K = D.read_opencv_matrix(p, "K")
    assert K.dtype == np.float32
    np.testing.assert_array_equal(K, np.array([[1, 2.5], [-3, 0.4]], np.float32))
    with pytest.raises(KeyError):
        D.read_opencv_matrix(p, "R")


def test_bgr_to_gray_known_answers():
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]], np.uint8)  # B, G, R
    np.testing.assert_array_equal(D.bgr_to_gray(px), [[29, 150, 76, 255, 0]])
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (64, 64, 3), dtype=np.uint8)
    ref = 0.114 * img[..., 0] + 0.587 * img[..., 1] + 0.299 * img[..., 2]
    assert np.max(np.abs(D.bgr_to_gray(img).astype(float) - ref)) <= 0.51


def test_equalize_hist():
    np.testing.assert_array_equal(D.equalize_hist(np.array([[0, 1, 2, 3]], np.uint8)), [[0, 85, 170, 255]])
    np.testing.assert_array_equal(D.equalize_hist(np.array([[7, 7], [9, 9]], np.uint8)), [[0, 0], [255, 255]])
    const = np.full((5, 6), 42, np.uint8)
    np.testing.assert_array_equal(D.equalize_hist(const), const)
    rng = np.random.default_rng(1)
    img = np.clip(rng.normal(120, 20, (80, 90)), 0, 255).astype(np.uint8)
    eq = D.equalize_hist(img)
    assert eq[img == img.min()].max() == 0 and eq[img == img.max()].min() == 255
    order = np.argsort(img.ravel(), kind="stable")
    assert np.all(np.diff(eq.ravel()[order].astype(int)) >= 0)  # monotone in the input
    # a flattened histogram: each output level's share of pixels approaches uniform
    cdf = np.cumsum(np.bincount(eq.ravel(), minlength=256)) / eq.size
    assert np.max(np.abs(cdf - np.arange(1, 257) / 256)) < 0.05


def test_circle_mask():
    m = D.circle_mask((31, 40), 12)
    assert m.dtype == np.uint8 and set(np.unique(m)) == {0, 255}
    brute = sum(1 for y in range(31) for x in range(40) if (x - 20) ** 2 + (y - 15) ** 2 <= 144)
    assert np.count_nonzero(m) == brute
    assert m[15, 20] == 255 and m[15, 32] == 255 and m[15, 33] == 0
