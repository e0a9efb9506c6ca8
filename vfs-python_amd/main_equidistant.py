"""Depth of one ICRA dataset frame (Python/main_equidistant.py:39-96), without OpenCV
or Open3D.

    python vfs-python_amd/main_equidistant.py /path/to/icra_dataset --id 0 --out out/

Writes, as the reference does: the disc mask (icra_dataset_mask.png), both views
(img_ref.png, img_other.png), the ground-truth point cloud coloured by the reference
view (icra_dataset_NNNN.pcd), and the colour-mapped solved and ground-truth depths
(depth.png, depth_gt.png).  Prints the median relative depth error against the
dataset's range.  The reference's interactive point-cloud viewer is not part of it.
"""

import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import colormap as C  # noqa: E402
import point_cloud as PC  # noqa: E402
import vfs_icra_dataset as D  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("dataset")
    ap.add_argument("--id", type=int, default=0)
    ap.add_argument("--near", type=float, default=1.5)
    ap.add_argument("--far", type=float, default=10.0)
    ap.add_argument("--out", default=".")
    ap.add_argument("--stored-vector-fields", action="store_true",
                    help="use the dataset's .flo vector fields instead of regenerating them from the pose")
    ap.add_argument("--solver", default="blocked", choices=["graph", "eager", "python", "blocked", "blocked_eager"])
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    ld = D.dataset_loader(a.dataset)
    r = D.depth_eval(ld, a.id, regenerate_vector_field=not a.stored_vector_fields, solver=a.solver)
    out = lambda name: os.path.join(a.out, name)  # noqa: E731
    C.write_bgr_png(out("icra_dataset_mask.png"), r["mask"])
    C.write_bgr_png(out("img_ref.png"), r["img_ref"])
    C.write_bgr_png(out("img_other.png"), r["img_other"])
    pts, cols = PC.colored_points(r["xyz_ref"], r["img_ref"])
    PC.write_pcd(out(f"icra_dataset_{a.id:04d}.pcd"), pts, cols)
    C.write_bgr_png(out("depth.png"), C.depth_to_colormap(r["z"], a.near, a.far))
    C.write_bgr_png(out("depth_gt.png"), C.depth_to_colormap(r["z_ref"], a.near, a.far))
    print(f"frame {a.id}: median relative depth error {r['err']:.4f}")
    return r


if __name__ == "__main__":
    main()
