"""Depth colouring of Python/colormap.py:25-292.

``colormap`` is the 256-entry jet table in B, G, R order (as cv2.imwrite expects),
built from jet's definition rather than stored: three trapezoids (red, green, blue)
of ramp length n = ceil(m / 4), each offset by n from the next, rounded to four
decimals like the reference's printed table.  ``depth_to_colormap`` maps inverse
depth linearly between 1/max_z and 1/min_z onto the table.
"""

import numpy as np


def jet(m=256):
    """(m, 3) jet table, R, G, B in [0, 1], unrounded."""
    n = -(-m // 4)
    u = np.concatenate([np.arange(1, n + 1) / n, np.ones(n - 1), np.arange(n, 0, -1) / n])
    g = -(-n // 2) - (1 if m % 4 == 1 else 0) + np.arange(1, u.size + 1)  # 1-based rows of the green ramp
    table = np.zeros((m, 3))
    for col, rows in ((0, g + n), (1, g), (2, g - n)):
        ok = (rows >= 1) & (rows <= m)
        table[rows[ok] - 1, col] = u[ok]
    return table


colormap_raw = np.round(jet(256), 4)
colormap = np.fliplr(colormap_raw)


def depth_to_colormap(depth, min_z, max_z):
    """Python/colormap.py:286-292: (H, W, 3) uint8 B, G, R; black where depth <= 0.

    index = round(clip(1/(depth + 1e-6) - 1/max_z, 0, 1/min_z - 1/max_z)
                  / (1/min_z - 1/max_z) * 255),  colour = floor(colormap[index] * 255).
    Non-finite depths are black too (the reference's integer cast of NaN has no
    defined index).  Accepts NumPy arrays or torch tensors.
    """
    if hasattr(depth, "detach"):
        depth = depth.detach().cpu().numpy()
    depth = np.asarray(depth)
    mask = np.isfinite(depth) & (depth > 0)
    d = np.where(mask, depth, 1.0)
    delta = 1 / min_z - 1 / max_z
    index = np.round(np.maximum(0, np.minimum(1 / (d + 1e-6) - 1 / max_z, delta) / delta) * 255).astype(np.int64)
    img = (colormap[index] * 255).astype(np.uint8)
    img[~mask] = (0, 0, 0)
    return img


def write_bgr_png(path, img):
    """cv2.imwrite of an 8-bit image: (H, W, 3) B, G, R or (H, W) gray."""
    from PIL import Image
    img = np.asarray(img, dtype=np.uint8)
    Image.fromarray(img[..., ::-1].copy() if img.ndim == 3 else img).save(path)
