import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vfs-python_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libvfs_hip.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_case():
    """The 256x256 plain workload: seeded pair, float32 vector fields, params (oracle fields)."""
    from oracle import vfs_oracle as O
    import synthetic as S
    pair = S.workload_pair("plain_small_256", seed=0)
    Hh, Ww = pair["img_ref"].shape
    cam = pair["cam"]
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((Hh, Ww)), cam, cam, pair["R"], pair["t"], 1)
    tvf = S.sanitize_field(tvf)
    cvf = np.asarray(cvf, dtype=np.float32)
    prm = O.OracleParams(Hh, Ww)
    S.apply_workload_params(prm, "plain_small_256")
    return dict(pair=pair, tvf=tvf, cvf=cvf, prm=prm)


@pytest.fixture(scope="session")
def small_oracle(small_case):
    """float64 and float32 oracle solves of the 256x256 workload, with per-level traces."""
    from oracle import vfs_oracle as O
    c = small_case
    args = (c["pair"]["img_ref"], c["pair"]["img_other"], c["pair"]["mask"], c["tvf"], c["cvf"], c["prm"])
    tr64, tr32 = {}, {}
    f64 = O.solve_stereo(*args, trace=tr64)
    with O.precision(np.float32):
        f32 = O.solve_stereo(*args, trace=tr32)
    return dict(flow=f64, trace=tr64, flow32=f32, trace32=tr32)


def flow_diff_stats(flow, ref, mask):
    d = np.linalg.norm(np.asarray(flow, np.float64) - np.asarray(ref, np.float64), axis=-1)[np.asarray(mask) > 0]
    return dict(median=float(np.median(d)), p90=float(np.percentile(d, 90)), p99=float(np.percentile(d, 99)),
                max=float(d.max()))
