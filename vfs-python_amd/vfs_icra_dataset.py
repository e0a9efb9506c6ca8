"""Reader of the ICRA fisheye stereo dataset layout (Python/vfs_icra_dataset.py:30-89)
without OpenCV, and the image preparation of Python/main_equidistant.py:40-80.

Layout under ``base_path`` (file stem shared by all parts of one frame):

    calib/<stem>.xml                         OpenCV FileStorage XML: R (3x3), t (3x1)
    image_02/data/<stem>.png, image_03/...    the two fisheye views (8-bit)
    proj_depth/groundtruth/image_02/<stem>.png   range (16-bit PNG, metres x 256)
    result_ours/<stem>depth.png              a published result (16-bit PNG, x 256)
    translationVector/<stem>.flo, calibrationVector/<stem>.flo   vector fields
                                             (Middlebury .flo)

``load_image`` returns B, G, R channel order as cv2.imread does.
"""

import glob
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

FLO_MAGIC = 202021.25  # the float whose bytes read "PIEH"


def read_flo(path):
    """(H, W, 2) float32 of a Middlebury .flo file (cv2.readOpticalFlow)."""
    with open(path, "rb") as f:
        magic = struct.unpack("<f", f.read(4))[0]
        if magic != FLO_MAGIC:
            raise ValueError(f"{path}: not a .flo file (magic {magic})")
        w, h = struct.unpack("<ii", f.read(8))
        if w <= 0 or h <= 0 or w * h > (1 << 30):
            raise ValueError(f"{path}: bad size {w} x {h}")
        data = np.frombuffer(f.read(4 * 2 * w * h), dtype="<f4")
    if data.size != 2 * w * h:
        raise ValueError(f"{path}: truncated")
    return data.reshape(h, w, 2).astype(np.float32)


def write_flo(path, flow):
    """Middlebury .flo writer (cv2.writeOpticalFlow)."""
    flow = np.asarray(flow, dtype="<f4")
    h, w = flow.shape[:2]
    with open(path, "wb") as f:
        f.write(struct.pack("<fii", FLO_MAGIC, w, h))
        f.write(np.ascontiguousarray(flow).tobytes())


_DT = {"d": np.float64, "f": np.float32, "i": np.int32, "u": np.uint8, "c": np.int8, "w": np.uint16,
       "s": np.int16}


def read_opencv_matrix(path, name):
    """Matrix `name` of an OpenCV FileStorage XML file (type_id "opencv-matrix")."""
    root = ET.parse(path).getroot()
    node = root.find(name)
    if node is None:
        raise KeyError(f"{path}: no node {name!r}")
    rows, cols = int(node.findtext("rows")), int(node.findtext("cols"))
    dt = node.findtext("dt").strip()
    vals = np.array(node.findtext("data").split(), dtype=np.float64)
    if vals.size != rows * cols:
        raise ValueError(f"{path}: {name} has {vals.size} values for {rows} x {cols}")
    return vals.reshape(rows, cols).astype(_DT.get(dt[-1], np.float64))


def _read_png(path, unchanged):
    from PIL import Image
    img = Image.open(path)
    if unchanged:
        return np.array(img)
    a = np.array(img.convert("RGB"))
    return a[..., ::-1].copy()  # B, G, R as cv2.imread


class dataset_loader():
    """Python/vfs_icra_dataset.py:30-89."""

    def __init__(self, base_path):
        self.base_path = base_path
        self.pose_path = os.path.join(base_path, "calib")
        self.fn_xmls = sorted(glob.glob(os.path.join(self.pose_path, "*.xml")))
        self.depth_scale = 256  # range in metres x 256

    def get_calibration(self):
        # equidistant model of the dataset: f = 800 / (2 pi / 3), centre (400, 400)
        return 800.0 / (2 * np.pi / 3), 400, 400

    def _path(self, i, sub, suffix=".png"):
        stem = os.path.splitext(os.path.basename(self.fn_xmls[i]))[0]
        return os.path.join(self.base_path, sub, stem + suffix)

    def load_pose(self, i):
        return read_opencv_matrix(self.fn_xmls[i], "R"), read_opencv_matrix(self.fn_xmls[i], "t")

    def load_depth(self, i):
        d = _read_png(self._path(i, os.path.join("proj_depth", "groundtruth", "image_02")), True)
        return (d / self.depth_scale).astype(np.float32)

    def load_result(self, i):
        d = _read_png(self._path(i, "result_ours", "depth.png"), True)
        return (d / self.depth_scale).astype(np.float32)

    def load_image(self, i):
        return (_read_png(self._path(i, os.path.join("image_02", "data")), False),
                _read_png(self._path(i, os.path.join("image_03", "data")), False))

    def load_vector_field(self, i):
        return (read_flo(self._path(i, "translationVector", ".flo")),
                read_flo(self._path(i, "calibrationVector", ".flo")))


def bgr_to_gray(img):
    """cv2.cvtColor(img, COLOR_BGR2GRAY) for uint8: fixed-point luma with 14-bit
    coefficients (0.114, 0.587, 0.299) and rounding."""
    img = np.asarray(img, dtype=np.int64)
    y = (img[..., 0] * 1868 + img[..., 1] * 9617 + img[..., 2] * 4899 + (1 << 13)) >> 14
    return y.astype(np.uint8)


def equalize_hist(gray):
    """cv2.equalizeHist for uint8: lut[j] = round(255 * (cdf[j] - cdf[first]) / (N - cdf[first]))
    from the first non-empty bin on, 0 below it; a constant image maps to itself."""
    g = np.asarray(gray, dtype=np.uint8)
    hist = np.bincount(g.ravel(), minlength=256)
    first = int(np.flatnonzero(hist)[0])
    total = g.size
    if hist[first] == total:
        return g.copy()
    scale = np.float32(255.0) / np.float32(total - hist[first])
    cum = np.cumsum(hist).astype(np.float64) - hist[first]
    lut = np.zeros(256, dtype=np.uint8)
    vals = np.rint(cum[first + 1:].astype(np.float32) * scale)
    lut[first + 1:] = np.clip(vals, 0, 255).astype(np.uint8)
    return lut[g]


def circle_mask(shape, radius, margin=0):
    """uint8 {0, 255} disc of `radius` px about (W // 2, H // 2) (the mask of
    Python/main_equidistant.py:52: radius W // 2 - 10), by the integer test
    (x - cx)^2 + (y - cy)^2 <= r^2."""
    h, w = shape[:2]
    cy, cx = h // 2, w // 2
    Y, X = np.mgrid[0:h, 0:w]
    m = ((X - cx) ** 2 + (Y - cy) ** 2) <= radius * radius
    if margin:
        m[:margin] = m[-margin:] = False
        m[:, :margin] = m[:, -margin:] = False
    return m.astype(np.uint8) * 255


def depth_eval(loader, i, param=None, regenerate_vector_field=True, solver="blocked"):
    """Python/main_equidistant.py:44-80 on frame i: equalised grayscale views, the disc
    mask, vector fields from the pose, solve, triangulate.

    Returns a dict: ``xyz`` and ``xyz_ref`` (H, W, 3) NumPy, the solved points and the
    dataset's range turned into points; ``z`` and ``z_ref`` their depths; ``mask``,
    ``img_ref`` (B, G, R); ``err``, the median relative depth error over the pixels
    where both depths are valid."""
    import torch
    import vfs as V
    import vfs_params as P
    from equidistant_cam import EquidistantCam
    from vector_field import get_grid_zero, get_vector_fields, unit_normalizer
    cam = EquidistantCam(loader.get_calibration())
    img_ref, img_other = loader.load_image(i)
    R, t = loader.load_pose(i)
    D_ref = loader.load_depth(i)
    mask = circle_mask(img_ref.shape, img_ref.shape[1] // 2 - 10)
    cam.setup_module()
    cam.set_mask(mask)
    if regenerate_vector_field:
        tvf, cvf = get_vector_fields(get_grid_zero(img_ref.shape[:2]), cam, cam, R, t, 1, unit_normalizer)
    else:
        tvf, cvf = loader.load_vector_field(i)
    # the ray -t/|t| on the image rim gives 0/0 in the unit normalisation; sampling
    # would spread that NaN, so such entries carry 0
    tvf = np.where(np.isfinite(tvf), tvf, 0).astype(np.float32)
    equi_ref, equi_other = equalize_hist(bgr_to_gray(img_ref)), equalize_hist(bgr_to_gray(img_other))
    if param is None:
        param = P.vfs_params(equi_ref)
    v = V.VFS(param)
    v.solver = solver
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(mask)
    v.create_vector_field_pyramid(tvf, cvf)
    v.solve_stereo(equi_ref, equi_other)
    xyz = v.triangulate_flow(cam, t).cpu().numpy()
    xyz_ref = cam.get_xyz_from_distance(torch.as_tensor(D_ref, device="cuda")).cpu().numpy()
    z, z_ref = xyz[..., 2], xyz_ref[..., 2]
    ok = (mask > 0) & (z_ref > 1e-7) & np.isfinite(z) & (z > 0)
    err = float(np.median(np.abs(z[ok] - z_ref[ok]) / z_ref[ok])) if ok.any() else float("nan")
    return dict(xyz=xyz, xyz_ref=xyz_ref, z=z, z_ref=z_ref, mask=mask, img_ref=img_ref, img_other=img_other,
                err=err)
