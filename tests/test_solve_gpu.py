"""End-to-end parity of VFS.solve_stereo / triangulate_flow on the GPU against the oracle.

Workload: the seeded 256x256 synthetic fisheye pair (f = 256/pi, t = (0.2, 0, 0) m,
1 deg rotation), 3 levels x 10 warps x 20 primal-dual iterations, lambda = 50,
alpha0 = 5, alpha1 = 1, beta = 10, gamma = 0.5.

Tolerances (fp32 GPU against the float64 oracle after the full iteration count):
  * flow: median |d flow| <= 1e-4 px over the mask; 99th percentile <= 1e-3 px;
  * depth: median relative depth difference <= 1e-3 over the mask;
  * the final mask (flow zero outside it) is exact.
"""

import numpy as np
import pytest

from oracle import vfs_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def oracle_flow(small_case):
    c = small_case
    return O.solve_stereo(c["pair"]["img_ref"], c["pair"]["img_other"], c["pair"]["mask"], c["tvf"], c["cvf"], c["prm"])


def _solver(small_case, mode):
    import vfs as V
    import vfs_params as P
    import synthetic as S
    pair = small_case["pair"]
    prm = P.vfs_params(pair["img_ref"])
    S.apply_workload_params(prm, "plain_small_256")
    v = V.VFS(prm)
    v.solver = mode
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pair["mask"])
    v.create_vector_field_pyramid(small_case["tvf"], small_case["cvf"])
    return v


def _flow_stats(flow, ref, mask):
    d = np.linalg.norm(flow.astype(np.float64) - ref, axis=-1)[mask > 0]
    return np.median(d), np.percentile(d, 99), d.max()


def test_solve_graph_matches_oracle(T, small_case, oracle_flow):
    v = _solver(small_case, "graph")
    flow = v.solve_stereo(small_case["pair"]["img_ref"], small_case["pair"]["img_other"])
    T.cuda.synchronize()
    f = flow.cpu().numpy()
    mask = small_case["pair"]["mask"]
    med, p99, mx = _flow_stats(f, oracle_flow, mask)
    print(f"flow |d| vs oracle: median {med:.3e} p99 {p99:.3e} max {mx:.3e} px")
    assert np.all(f[mask == 0] == 0)
    assert med <= 1e-4
    assert p99 <= 1e-3


def test_solver_modes_agree_bitwise(T, small_case):
    pair = small_case["pair"]
    outs = {}
    for mode in ("graph", "eager", "python"):
        v = _solver(small_case, mode)
        outs[mode] = v.solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
        if mode == "graph":  # replay of the recorded graph gives the same result
            again = v.solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
            np.testing.assert_array_equal(again, outs[mode])
    np.testing.assert_array_equal(outs["graph"], outs["eager"])
    np.testing.assert_array_equal(outs["graph"], outs["python"])


def test_depth_matches_oracle_and_ground_truth(T, small_case, oracle_flow):
    import equidistant_cam as E
    pair = small_case["pair"]
    v = _solver(small_case, "graph")
    v.solve_stereo(pair["img_ref"], pair["img_other"])
    cam = E.EquidistantCam(pair["cam"])
    cam.setup_module()
    xyz = v.triangulate_flow(cam, pair["t"]).cpu().numpy()
    xyz_o = O.triangulate_flow(oracle_flow, pair["cam"], pair["t"].reshape(3), pair["mask"])
    m = pair["mask"] > 0
    z, zo = xyz[..., 2][m], xyz_o[..., 2][m]
    ok = np.isfinite(zo) & (np.abs(zo) > 1e-3)
    rel = np.abs(z[ok] - zo[ok]) / np.abs(zo[ok])
    print(f"depth rel diff vs oracle: median {np.median(rel):.3e} p99 {np.percentile(rel, 99):.3e}")
    assert np.median(rel) <= 1e-3
    assert np.all(xyz[~m] == [0, 0, -1])
