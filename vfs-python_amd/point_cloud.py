"""Point-cloud export of Python/main_equidistant.py:80-88 without Open3D.

The reference hands (H*W, 3) points and (H*W, 3) R, G, B colours in [0, 1] to
Open3D, which writes a PCD file.  ``write_pcd`` writes PCD v0.7 itself: fields
x y z, plus rgb as 0x00RRGGBB whose four bytes are declared a float (TYPE F, the
PCL convention that Open3D and PCL read), in binary or ASCII; ``read_pcd`` reads
both back, with rgb declared F or U.
"""

import numpy as np


def colored_points(xyz, img_bgr):
    """points (N, 3) float32 and colours (N, 3) R, G, B in [0, 1] of a (H, W, 3) xyz map
    and its (H, W, 3) uint8 B, G, R view (Python/main_equidistant.py:85-87)."""
    if hasattr(xyz, "detach"):
        xyz = xyz.detach().cpu().numpy()
    pts = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
    cols = np.fliplr(np.asarray(img_bgr).reshape(-1, 3) / 255)
    return pts, cols


def _pack_rgb(colors):
    c = np.clip(np.rint(np.asarray(colors, dtype=np.float64) * 255), 0, 255).astype(np.uint32)
    return (c[:, 0] << 16) | (c[:, 1] << 8) | c[:, 2]


def write_pcd(path, points, colors=None, binary=True):
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
    n = pts.shape[0]
    fields, size, typ = "x y z", "4 4 4", "F F F"
    if colors is not None:
        fields, size, typ = fields + " rgb", size + " 4", typ + " F"
    cnt = " ".join(["1"] * len(fields.split()))
    header = (f"# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS {fields}\nSIZE {size}\nTYPE {typ}\n"
              f"COUNT {cnt}\nWIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\n"
              f"DATA {'binary' if binary else 'ascii'}\n")
    if colors is not None:
        rec = np.empty(n, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("rgb", "<u4")])
        rec["rgb"] = _pack_rgb(np.asarray(colors).reshape(-1, 3))
    else:
        rec = np.empty(n, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
    rec["x"], rec["y"], rec["z"] = pts[:, 0], pts[:, 1], pts[:, 2]
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        if binary:
            f.write(rec.tobytes())
        else:
            for r in rec:
                vals = [repr(float(r[k])) for k in ("x", "y", "z")]
                if colors is not None:
                    vals.append(repr(float(np.uint32(r["rgb"]).view(np.float32))))
                f.write((" ".join(vals) + "\n").encode("ascii"))


def read_pcd(path):
    """points (N, 3) float32 and colours (N, 3) in [0, 1] (None without an rgb field)."""
    with open(path, "rb") as f:
        head = {}
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: no DATA line")
            line = line.decode("ascii").strip()
            if not line or line.startswith("#"):
                continue
            key, _, val = line.partition(" ")
            head[key] = val
            if key == "DATA":
                break
        fields = head["FIELDS"].split()
        n = int(head["POINTS"])
        kinds = dict(F="<f", U="<u", I="<i")
        dt = [(nm, kinds[t] + s) for nm, t, s in zip(fields, head["TYPE"].split(), head["SIZE"].split())]
        if head["DATA"] == "binary":
            rec = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt)
        elif head["DATA"] == "ascii":
            rows = [ln.split() for ln in f.read().decode("ascii").splitlines() if ln.strip()]
            rec = np.empty(n, dtype=dt)
            for j, (nm, t) in enumerate(dt):
                rec[nm] = np.array([r[j] for r in rows], dtype=np.float64 if t[1] == "f" else np.int64)
        else:
            raise ValueError(f"{path}: unsupported DATA {head['DATA']}")
    if rec.shape[0] != n:
        raise ValueError(f"{path}: truncated")
    pts = np.stack([rec["x"], rec["y"], rec["z"]], axis=1).astype(np.float32)
    cols = None
    if "rgb" in fields:
        v = np.ascontiguousarray(rec["rgb"])
        v = v.view("<u4") if v.dtype.kind == "f" else v.astype(np.uint32)
        cols = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], axis=1) / 255.0
    return pts, cols
