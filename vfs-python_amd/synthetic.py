"""Seeded synthetic fisheye stereo pairs with known depth.

Scene model (one pair):

* equidistant camera r = f*theta with f = min(H, W) / pi (180 deg FOV), principal
  point (W/2, H/2) (the convention of Python/vfs_icra_dataset.py:42-44);
* range field D (metres along the ray) = base + Gaussian bumps + one planar
  step, clipped; the z-depth that triangulation returns is D * cos(theta);
* texture: multi-octave value noise, quantised to uint8;
* second view: a point P seen by the reference camera is seen by the other
  camera at R (P + t); the other image is rendered by inverse warping through
  the true range field (fixed-point inversion of the forward map), plus
  Gaussian noise, then quantised to uint8;
* mask: the 180 deg image circle with an outer ring of ``margin`` pixels forced
  to zero, so that every pyramid level keeps a zero border ring.

This is input generation for tests and benchmarks.  All random draws come from
NumPy's PCG64 on the host; the per-pixel evaluation runs in float64 NumPy, or
in float64 torch on a GPU when ``device`` is given (same formulas; the two may
differ in the last bits of a transcendental and hence, after quantisation, by
one grey level at isolated pixels).
"""

import math

import numpy as np


class _NumpyOps:
    hypot = staticmethod(np.hypot)
    atan2 = staticmethod(np.arctan2)
    cos = staticmethod(np.cos)
    sin = staticmethod(np.sin)
    floor = staticmethod(np.floor)
    exp = staticmethod(np.exp)
    where = staticmethod(np.where)

    @staticmethod
    def stack(a, axis):
        return np.stack(a, axis)

    @staticmethod
    def clip(a, lo, hi):
        return np.clip(a, lo, hi)

    @staticmethod
    def index_mod(a, n):
        return np.mod(a.astype(np.int64), n)

    @staticmethod
    def table(t):
        return t

    @staticmethod
    def full(shape, v):
        return np.full(shape, v, dtype=np.float64)

    @staticmethod
    def grid(H, W):
        Y, X = np.mgrid[0:H, 0:W].astype(np.float64)
        return X, Y

    @staticmethod
    def matT(R):
        return np.asarray(R, dtype=np.float64).T

    @staticmethod
    def host(a):
        return a


class _TorchOps:
    def __init__(self, device):
        import torch
        self.torch = torch
        self.device = device
        self.hypot = torch.hypot
        self.atan2 = torch.atan2
        self.cos = torch.cos
        self.sin = torch.sin
        self.floor = torch.floor
        self.exp = torch.exp

    def where(self, c, a, b):
        t = self.torch
        a = a if isinstance(a, t.Tensor) else t.tensor(a, dtype=t.float64, device=self.device)
        b = b if isinstance(b, t.Tensor) else t.tensor(b, dtype=t.float64, device=self.device)
        return t.where(c, a, b)

    def stack(self, a, axis):
        return self.torch.stack(a, axis)

    def clip(self, a, lo, hi):
        return self.torch.clamp(a, lo, hi)

    def index_mod(self, a, n):
        return self.torch.remainder(a.long(), n)

    def table(self, t):
        return self.torch.as_tensor(t, dtype=self.torch.float64, device=self.device)

    def full(self, shape, v):
        return self.torch.full(shape, v, dtype=self.torch.float64, device=self.device)

    def grid(self, H, W):
        t = self.torch
        Y, X = t.meshgrid(t.arange(H, dtype=t.float64, device=self.device),
                          t.arange(W, dtype=t.float64, device=self.device), indexing="ij")
        return X, Y

    def matT(self, R):
        return self.torch.as_tensor(np.asarray(R, dtype=np.float64).T, device=self.device)

    def host(self, a):
        return a.cpu().numpy()


def _project(xp, xyz, f, cx, cy):
    r = xp.hypot(xyz[..., 0], xyz[..., 1])
    th = xp.atan2(r, xyz[..., 2])
    ph = xp.atan2(xyz[..., 1], xyz[..., 0])
    return f * th * xp.cos(ph) + cx, f * th * xp.sin(ph) + cy


def _unproject(xp, x, y, f, cx, cy):
    dx = x - cx
    dy = y - cy
    th = xp.hypot(dx, dy) / f
    ph = xp.atan2(dy, dx)
    st = xp.sin(th)
    return xp.stack([st * xp.cos(ph), st * xp.sin(ph), xp.cos(th)], -1)


def rotation_about_axis(axis, angle_rad):
    """Rodrigues rotation matrix."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle_rad) * K + (1 - math.cos(angle_rad)) * (K @ K)


class ValueNoise:
    """Multi-octave value noise evaluated at continuous pixel coordinates."""

    def __init__(self, rng, base_cell, octaves=5, lattice=256):
        self.cells = [base_cell / (2 ** o) for o in range(octaves)]
        self.amps = [0.5 ** o for o in range(octaves)]
        self.tables = [rng.random((lattice, lattice)) for _ in range(octaves)]
        self.lattice = lattice

    def __call__(self, x, y, xp=_NumpyOps):
        acc = 0.0
        n = self.lattice
        for cell, amp, tab in zip(self.cells, self.amps, self.tables):
            tab = xp.table(tab)
            gx = x / cell
            gy = y / cell
            x0 = xp.floor(gx)
            y0 = xp.floor(gy)
            fx = gx - x0
            fy = gy - y0
            sx = fx * fx * (3 - 2 * fx)
            sy = fy * fy * (3 - 2 * fy)
            ix = xp.index_mod(x0, n)
            iy = xp.index_mod(y0, n)
            ix1 = (ix + 1) % n
            iy1 = (iy + 1) % n
            top = tab[iy, ix] * (1 - sx) + tab[iy, ix1] * sx
            bot = tab[iy1, ix] * (1 - sx) + tab[iy1, ix1] * sx
            acc = acc + amp * (top * (1 - sy) + bot * sy)
        return acc / sum(self.amps)


class RangeField:
    """D(x, y) = base + sum of Gaussian bumps + a planar step, clipped to [dmin, dmax]."""

    def __init__(self, rng, H, W, base=4.0, n_bumps=8, amp=1.5, sigma_px=(10.0, 40.0),
                 step=1.0, clip=(1.0, 10.0)):
        s = min(H, W) / 256.0
        self.base = base
        self.cx = rng.uniform(0, W, n_bumps)
        self.cy = rng.uniform(0, H, n_bumps)
        self.amp = rng.uniform(-amp, amp, n_bumps)
        self.sig = rng.uniform(sigma_px[0], sigma_px[1], n_bumps) * s
        ang = rng.uniform(0, 2 * math.pi)
        self.nx, self.ny = math.cos(ang), math.sin(ang)
        self.off = rng.uniform(-0.25, 0.25) * min(H, W)
        self.px, self.py = W / 2.0, H / 2.0
        self.step = step * (1 if rng.random() < 0.5 else -1)
        self.clip = clip

    def __call__(self, x, y, xp=_NumpyOps):
        d = self.base + 0.0 * x
        for cx, cy, a, sg in zip(self.cx, self.cy, self.amp, self.sig):
            d = d + a * xp.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * sg * sg))
        side = (x - self.px) * self.nx + (y - self.py) * self.ny - self.off
        d = d + xp.where(side > 0, self.step, 0.0)
        return xp.clip(d, *self.clip)


def circle_mask(H, W, margin, radius=None):
    cx, cy = W / 2.0, H / 2.0
    r = min(H, W) / 2.0 if radius is None else radius
    Y, X = np.mgrid[0:H, 0:W]
    m = ((X - cx) ** 2 + (Y - cy) ** 2) <= r * r
    if margin > 0:
        m[:margin, :] = False
        m[-margin:, :] = False
        m[:, :margin] = False
        m[:, -margin:] = False
    return (m.astype(np.uint8) * 255)


def make_pair(H, W, seed, t=(0.2, 0.0, 0.0), rot_deg=1.0, rot_random=False, n_levels=3,
              depth_base=4.0, depth_clip=(1.0, 10.0), contrast=1.0, noise_std=1.0 / 255.0,
              fp_iters=30, device=None, level_ratio=2.0, f=None, tex_cell=None):
    """Return a dict describing one seeded synthetic pair.

    Keys: img_ref, img_other (H, W) uint8; mask (H, W) uint8 in {0, 255}; range_gt, z_gt
    (H, W) float64; R (3, 3); t (3, 1); cam (f, cx, cy).  The focal length defaults to
    min(H, W) / pi (a 180 deg image circle).
    ``rot_random`` draws the rotation angle uniformly in [0, rot_deg].  ``device``
    (e.g. "cuda") evaluates the per-pixel math with torch on that device.
    """
    xp = _NumpyOps if device is None else _TorchOps(device)
    rng = np.random.Generator(np.random.PCG64(seed))
    f = min(H, W) / math.pi if f is None else float(f)
    cx, cy = W / 2.0, H / 2.0
    axis = rng.normal(size=3)
    ang = math.radians(rng.uniform(0, rot_deg) if rot_random else rot_deg)
    R = rotation_about_axis(axis, ang)
    tv = np.asarray(t, dtype=np.float64).reshape(3)
    depth = RangeField(rng, H, W, base=depth_base, clip=depth_clip)
    tex = ValueNoise(rng, base_cell=max(8.0, min(H, W) / 8.0) if tex_cell is None else float(tex_cell))
    margin = int(math.ceil(level_ratio ** (n_levels - 1))) + 1  # one coarsest-level pixel

    X, Y = xp.grid(H, W)
    lo = 0.5 - 0.5 * contrast
    RT = xp.matT(R)
    tvx = xp.table(tv) if device is not None else tv

    def shade(x, y):
        return lo + contrast * tex(x, y, xp)

    i0 = shade(X, Y)

    def forward(x, y):
        d = _unproject(xp, x, y, f, cx, cy)
        P = depth(x, y, xp)[..., None] * d
        Q = (P + tvx) @ RT
        return _project(xp, Q, f, cx, cy)

    # inverse warp: for each pixel x' of the other view solve forward(x) = x'
    xs, ys = X + 0.0, Y + 0.0
    for _ in range(fp_iters):
        fx, fy = forward(xs, ys)
        xs = xs + (X - fx)
        ys = ys + (Y - fy)
    i1 = shade(xs, ys)
    rng_gt = depth(X, Y, xp)

    def quant(img):
        img = xp.host(img) + rng.normal(0.0, noise_std, size=(H, W))
        return np.clip(np.rint(img * 255.0), 0, 255).astype(np.uint8)

    img_ref = quant(i0)
    img_other = quant(i1)
    rng_gt = xp.host(rng_gt)
    Xh, Yh = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    theta = np.hypot(Xh - cx, Yh - cy) / f
    return dict(img_ref=img_ref, img_other=img_other,
                mask=circle_mask(H, W, margin), range_gt=rng_gt, z_gt=rng_gt * np.cos(theta),
                R=R, t=tv.reshape(3, 1), cam=(f, cx, cy))


# Named workloads (sizes and solver schedules) used by tests and the benchmark.
# The 256 x 256 workload states its solver parameters (lambda 50, alpha0 5, alpha1 1,
# a = beta 10, b = gamma 0.5).  The 800 x 800 and batched workloads state only the
# schedule; they use the reference's default parameters (Python/vfs_params.py:33-43).
# lambda 50 loses the 800 x 800 scene (median relative depth error 0.53-0.56 against
# the ground truth on seeds 0 and 1, the float64 oracle alike): the linearised data
# term's steps overshoot, and only increases of u are clamped.  The defaults solve it
# (0.033), as do lambda 10-20 with the other values kept (0.020-0.029).
_SOLVER = dict(Lambda=50.0, alpha0=5.0, alpha1=1.0, beta=10.0, gamma=0.5)
_REF_DEFAULTS = dict(Lambda=5.0, alpha0=17.0, alpha1=1.2, beta=9.0, gamma=0.85)
WORKLOADS = {
    "plain_small_256": dict(H=256, W=256, t=(0.2, 0.0, 0.0), rot_deg=1.0, rot_random=False,
                            nLevel=3, nWarpIters=10, nSolverIters=20, **_SOLVER),
    "plain_full_800": dict(H=800, W=800, t=(0.3, 0.0, 0.0), rot_deg=2.0, rot_random=True,
                           nLevel=5, nWarpIters=20, nSolverIters=50, **_REF_DEFAULTS),
    # gradient of the depth loss: 3 levels x 10 warps x 50 iterations, checkpoint every 25
    "adjoint_800": dict(H=800, W=800, t=(0.3, 0.0, 0.0), rot_deg=2.0, rot_random=True,
                        nLevel=3, nWarpIters=10, nSolverIters=50, checkpoint=25, pairs=8, **_REF_DEFAULTS),
    # the reference's default parameters (Python/vfs_params.py:25-43): 11 levels at a
    # ratio of 1.2 (800x800 down to 127x127), 100 warps x 100 iterations per level
    "default_800": dict(H=800, W=800, t=(0.3, 0.0, 0.0), rot_deg=2.0, rot_random=True,
                        nLevel=11, fScale=1.2, nWarpIters=100, nSolverIters=100,
                        Lambda=5.0, alpha0=17.0, alpha1=1.2, beta=9.0, gamma=0.85),
}

# 64 pairs, 16 each of four sizes, every pair with its own pose and scene.
BATCH_64 = dict(sizes=[(512, 512)] * 16 + [(800, 800)] * 16 + [(1024, 1024)] * 16 + [(960, 1280)] * 16,
                rot_deg=3.0, t_norm=(0.1, 0.4), nLevel=5, nWarpIters=20, nSolverIters=50, **_REF_DEFAULTS)


def batch_pairs(spec=BATCH_64, base_seed=1000, device=None, n=None):
    """Seeded pairs of a batch workload: pair i uses seed base_seed + i, its own rotation
    (angle uniform in [0, rot_deg]) and a baseline of length uniform in t_norm, mostly
    along x (triangulation uses the x coordinate, Python/vfs.cu:728-733)."""
    sizes = spec["sizes"][:n] if n else spec["sizes"]
    out = []
    for i, (h, w) in enumerate(sizes):
        rng = np.random.Generator(np.random.PCG64(base_seed + 10_000 + i))
        d = np.array([1.0, rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)])
        t = d / np.linalg.norm(d) * rng.uniform(*spec["t_norm"])
        out.append(make_pair(h, w, base_seed + i, t=tuple(t), rot_deg=spec["rot_deg"], rot_random=True,
                             n_levels=spec["nLevel"], device=device))
    return out


def workload_pair(name, seed=0, device=None):
    w = WORKLOADS[name]
    return make_pair(w["H"], w["W"], seed, t=w["t"], rot_deg=w["rot_deg"], rot_random=w["rot_random"],
                     n_levels=w["nLevel"], device=device, level_ratio=w.get("fScale", 2.0))


def sanitize_field(tvf):
    """Zero the non-finite entries of a translation vector field.

    With a 180 deg FOV and a translation in the image plane, the ray -t/|t| lies
    on the image circle's rim: there (f(XYZ + t) - f(XYZ - t)) / 2 is the zero
    vector and the unit normalisation of Python/vector_field.py:46-48 yields
    0/0.  Bilinear sampling would spread that NaN into unmasked pixels through
    the pyramid and the warp taps, so inputs built here carry 0 instead.
    """
    tvf = np.array(tvf, dtype=np.float32, copy=True)
    tvf[~np.isfinite(tvf)] = 0.0
    return tvf


def apply_workload_params(param, name_or_spec):
    """Set the solver fields of a vfs_params-like object for a named workload (fScale 2
    unless the workload names another)."""
    w = WORKLOADS[name_or_spec] if isinstance(name_or_spec, str) else name_or_spec
    for k in ("nLevel", "nWarpIters", "nSolverIters", "Lambda", "alpha0", "alpha1", "beta", "gamma"):
        setattr(param, k, w[k])
    param.fScale = w.get("fScale", 2.0)
    return param
