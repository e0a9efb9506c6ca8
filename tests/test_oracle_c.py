"""The C restatement of the oracle (oracle/vfs_oracle.c) against the NumPy one.

Both are float64 restatements of the same kernels in the same operation order, so
their flows agree to rounding: median |d flow| <= 1e-12 px and max |d flow| <= 1e-7 px
after whole solves, on the 64 x 64 golden pair, the 256 x 256 workload, a pair at the
reference's default ratio 1.2 with odd level sizes, and 60 iterations per warp.
Measured maxima: 4.8e-14, 7.9e-9 (the 256 x 256 workload, whose solve amplifies
rounding: its float32 restatement is 0.2 px off the float64 one at the 99th
percentile), 2.8e-14 and 2.0e-15 px.
"""

import os

import numpy as np
import pytest

from oracle import vfs_oracle as O
from oracle import vfs_oracle_c as C

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def built():
    C.build()


def _case(h, w, seed, levels, ratio=2.0):
    import synthetic as S
    pr = S.make_pair(h, w, seed, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=levels, level_ratio=ratio)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((h, w)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    return pr, S.sanitize_field(tvf), cvf.astype(np.float32)


def _agree(pr, tvf, cvf, prm, median_k=5, mask=None):
    mask = pr["mask"] if mask is None else mask
    fo = O.solve_stereo(pr["img_ref"], pr["img_other"], mask, tvf, cvf, prm) if median_k == 5 else None
    fc = C.solve_stereo(pr["img_ref"], pr["img_other"], mask, tvf, cvf, prm, median_k=median_k)
    return fo, fc


def test_c_oracle_matches_numpy_oracle_on_the_golden_pair():
    g = np.load(os.path.join(HERE, "golden", "solve_64.npz"))
    nl, nw, ni = (int(x) for x in g["schedule"])
    prm = O.OracleParams(64, 64, nLevel=nl, fScale=2.0, nWarpIters=nw, nSolverIters=ni, alpha0=5.0, alpha1=1.0,
                         beta=10.0, gamma=0.5, Lambda=50.0, limitRange=0.1)
    fc = C.solve_stereo(g["img_ref"], g["img_other"], g["mask"], g["tvf"], g["cvf"], prm)
    d = np.abs(fc - g["flow"]).max()
    print("max |d flow| against the golden flow", d)
    assert d <= 1e-7 and np.median(np.abs(fc - g["flow"])) <= 1e-12
    assert np.all(fc[g["mask"] == 0] == 0)


@pytest.mark.parametrize("case", ["plain_small_256", "default_ratio", "deep_schedule"])
def test_c_oracle_matches_numpy_oracle(case):
    import synthetic as S
    if case == "plain_small_256":
        pr = S.workload_pair("plain_small_256", seed=0)
        tvf, cvf = O.get_vector_fields(O.get_grid_zero((256, 256)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
        tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
        prm = O.OracleParams(256, 256)
        S.apply_workload_params(prm, "plain_small_256")
    elif case == "default_ratio":
        pr, tvf, cvf = _case(61, 77, 4, 4, ratio=1.2)
        prm = O.OracleParams(61, 77, nLevel=4, nWarpIters=3, nSolverIters=8)
        assert prm.fScale == 1.2
    else:
        pr, tvf, cvf = _case(48, 40, 9, 2)
        prm = O.OracleParams(48, 40, nLevel=2, fScale=2.0, nWarpIters=4, nSolverIters=60, Lambda=20.0)
    fo, fc = _agree(pr, tvf, cvf, prm)
    d = np.abs(fc - fo).max()
    print(case, "max |d flow|", d, "max |flow|", np.abs(fo).max())
    assert np.abs(fo).max() > 0.1
    assert d <= 1e-7 and np.median(np.abs(fc - fo)) <= 1e-12


def test_c_oracle_rejects_bad_arguments():
    prm = O.OracleParams(8, 8, nLevel=1, nWarpIters=1, nSolverIters=1)
    z8 = np.zeros((8, 8), np.uint8)
    z2 = np.zeros((8, 8, 2), np.float32)
    with pytest.raises(RuntimeError):
        C.solve_stereo(z8, z8, z8, z2, z2, prm, median_k=4)
    prm.nLevel = 5  # 8 -> 6 -> 5 -> 4 -> 3 at ratio 1.2 is fine; 40 levels is not
    C.solve_stereo(z8, z8, z8, z2, z2, prm)
    prm.nLevel = 40
    with pytest.raises(RuntimeError):
        C.solve_stereo(z8, z8, z8, z2, z2, prm)
