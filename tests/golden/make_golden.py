"""Write the golden fixtures under tests/golden/ from the float64 oracle.

Fixtures (inputs and expected outputs only, compressed .npz):
  kernels_24x32.npz  seeded random inputs of every per-kernel function at 24x32
                     with a ragged mask, and the oracle's outputs;
  solve_64.npz       a seeded 64x64 synthetic fisheye pair, float32 vector fields,
                     and the oracle's flow after 2 levels x 3 warps x 10 iterations;
  solve_800.npz      the 800x800 workload (seed 0: 5 levels x 20 warps x 50
                     iterations, the reference's default solver parameters): the
                     uint8 pair, mask and pose, and, on every 4th row and column, the
                     float64 oracle's flow and depth, the ground-truth depth and the
                     float32 restatement's flow (the spread that single-precision
                     rounding alone gives the solve).  About 4 minutes of CPU.

Run from the repository root: python tests/golden/make_golden.py [--only solve_800]
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "vfs-python_amd"))
sys.path.insert(0, ROOT)

from oracle import vfs_oracle as O  # noqa: E402
import synthetic as S  # noqa: E402

PRM = dict(alpha0=5.0, alpha1=1.0, beta=10.0, gamma=0.5, Lambda=50.0, limitRange=0.1)


def kernel_inputs(seed=2024, Hh=24, Ww=32):
    rng = np.random.default_rng(seed)
    f32 = lambda a: np.asarray(a, np.float32)  # noqa: E731
    mask = S.circle_mask(Hh, Ww, 2, radius=11.0)
    mask[rng.random((Hh, Ww)) < 0.05] = 0  # ragged holes
    ang = rng.uniform(0, 2 * np.pi, (Hh, Ww))
    return dict(
        mask=mask,
        img=f32(rng.random((Hh, Ww))),
        img2=f32(rng.random((Hh, Ww))),
        tvf=f32(np.stack([np.cos(ang), np.sin(ang)], -1) * rng.uniform(0.4, 1.5, (Hh, Ww, 1))),
        flow=f32(rng.normal(size=(Hh, Ww, 2)) * 2.5),
        p=f32(rng.normal(size=(Hh, Ww, 2)) * 0.3),
        q=f32(rng.normal(size=(Hh, Ww, 4)) * 0.3),
        tp=f32(rng.normal(size=(Hh, Ww, 2)) * 0.3),
        u=f32(rng.normal(size=(Hh, Ww))),
        u_=f32(rng.normal(size=(Hh, Ww))),
        v_=f32(rng.normal(size=(Hh, Ww, 2)) * 0.1),
        deriv=f32(rng.normal(size=(Hh, Ww, 2)) * 0.2),
        uf=f32(rng.normal(size=(Hh, Ww))),
        sigma=np.float32(3.25), tau=np.float32(0.4), mu=np.float32(0.8),
    )


def kernel_outputs(a):
    out = {}
    out["gaussian"] = O.gaussian_masked(a["img"], a["mask"])
    g32 = out["gaussian"].astype(np.float32).astype(np.float64)
    out["abc"], out["eta"] = O.compute_tensor_eta(g32, a["mask"], PRM["beta"], PRM["gamma"], PRM["alpha0"],
                                                  PRM["alpha1"])
    out["warped"] = O.warping_image(a["img"], a["flow"])
    out["tv"] = O.find_warping_vector(a["flow"], a["tvf"])
    out["deriv_out"] = O.compute_derivatives(a["img2"], a["img"], a["tvf"])
    out["p_out"], out["tp_out"], out["q_out"] = O.update_dual_variables(
        a["p"], a["q"], float(a["sigma"]), a["u_"], a["v_"], out["abc"].astype(np.float32), a["mask"],
        PRM["alpha0"], PRM["alpha1"])
    eta32 = (np.abs(out["eta"]) + 1.0).astype(np.float32)
    out["eta_in"] = eta32
    out["u_th"] = O.l1_thresholding(float(a["tau"]), a["tp"], a["u"], a["u_"], a["deriv"], eta32, a["mask"],
                                    PRM["Lambda"])
    out["ubar_out"], out["vbar_out"] = O.update_primal_variables(
        float(a["tau"]), float(a["mu"]), a["p"], a["q"], a["u"], a["u_"], a["v_"], out["abc"].astype(np.float32),
        eta32, a["mask"], PRM["alpha0"], PRM["alpha1"])
    out["median"] = O.median_filter(a["u"], a["mask"])
    out["flow_out"], out["u0_out"] = O.compute_optical_flow_with_clamp(a["flow"], a["u"], a["uf"], a["tvf"],
                                                                       a["mask"], PRM["limitRange"])
    out["up1"] = O.upsampling(a["u"], (48, 64), 2.0)
    out["down"] = O.downsample_bilinear(a["img"], (12, 16))
    out["mask_down"] = O.downsample_mask(a["mask"], (12, 16))
    return out


def solve_case(seed=5, n=64):
    pair = S.make_pair(n, n, seed, t=(0.2, 0.0, 0.0), rot_deg=1.0, n_levels=2)
    cam = pair["cam"]
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((n, n)), cam, cam, pair["R"], pair["t"], 1)
    tvf = S.sanitize_field(tvf)
    cvf = cvf.astype(np.float32)
    prm = O.OracleParams(n, n, nLevel=2, fScale=2.0, nWarpIters=3, nSolverIters=10, **PRM)
    flow = O.solve_stereo(pair["img_ref"], pair["img_other"], pair["mask"], tvf, cvf, prm)
    return dict(img_ref=pair["img_ref"], img_other=pair["img_other"], mask=pair["mask"], tvf=tvf, cvf=cvf,
                cam=np.array(cam), t=pair["t"].reshape(3), flow=flow,
                schedule=np.array([2, 3, 10], np.int64))


SUB = 4  # stride of the stored rows and columns of the 800x800 results


def _solve_800(args):
    """One oracle solve of the 800x800 workload in float64 or float32 (run in a pool)."""
    dtype, pair = args
    w = S.WORKLOADS["plain_full_800"]
    cam = pair["cam"]
    hh, ww = pair["img_ref"].shape
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), cam, cam, pair["R"], pair["t"], 1)
    tvf = S.sanitize_field(tvf)
    cvf = cvf.astype(np.float32)
    prm = O.OracleParams(hh, ww)
    S.apply_workload_params(prm, w)
    if dtype == "f32":
        with O.precision(np.float32):
            return O.solve_stereo(pair["img_ref"], pair["img_other"], pair["mask"], tvf, cvf, prm)
    return O.solve_stereo(pair["img_ref"], pair["img_other"], pair["mask"], tvf, cvf, prm)


def solve_800(seed=0):
    from multiprocessing import Pool
    pair = S.workload_pair("plain_full_800", seed=seed)
    with Pool(2) as pool:
        f64, f32 = pool.map(_solve_800, [("f64", pair), ("f32", pair)])
    xyz = O.triangulate_flow(f64, pair["cam"], pair["t"].reshape(3), pair["mask"])
    m = pair["mask"] > 0
    z, zg = xyz[..., 2], pair["z_gt"]
    ok = m & np.isfinite(z) & (zg > 1e-3)
    err = float(np.median(np.abs(z[ok] - zg[ok]) / zg[ok]))
    print(f"solve_800: oracle median relative depth error against the ground truth {err:.4f}")
    sub = (slice(None, None, SUB), slice(None, None, SUB))
    return dict(img_ref=pair["img_ref"], img_other=pair["img_other"], mask=pair["mask"], cam=np.array(pair["cam"]),
                R=pair["R"], t=pair["t"].reshape(3), seed=np.int64(seed), sub=np.int64(SUB),
                flow64=f64[sub].astype(np.float32), flow32=f32[sub].astype(np.float32),
                z64=z[sub].astype(np.float32), z_gt=zg[sub].astype(np.float32),
                oracle_err_vs_gt=np.float64(err))


def main():
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    if only in (None, "kernels_24x32"):
        a = kernel_inputs()
        np.savez_compressed(os.path.join(HERE, "kernels_24x32.npz"), **a, **kernel_outputs(a))
    if only in (None, "solve_64"):
        np.savez_compressed(os.path.join(HERE, "solve_64.npz"), **solve_case())
    if only in (None, "solve_800"):
        np.savez_compressed(os.path.join(HERE, "solve_800.npz"), **solve_800())
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
