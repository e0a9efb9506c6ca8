"""The 5 x 5 median selection network of csrc/vfs_median.h, read from the header as
compiled.

* Exhaustive 0-1 check: a comparator network selects the k-th smallest of every input
  if and only if it does so for every 0/1 input (the 0-1 principle, Knuth, TAOCP vol. 3,
  section 5.3.4).  All 2^25 0/1 windows are run at once as bit planes
  (min = AND, max = OR) and output 12 must be 1 exactly when at least 13 inputs are 1.
* Float windows with ties and signed values against sorting (the median of 25 values is
  one of them, so the comparison is exact).
"""

import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vfs-python_amd", "csrc",
                   "vfs_median.h")


def _network():
    src = open(HDR).read()
    body = src[src.index("median_select<5>(float (&v)[25])"):src.index("return v[12];")]
    ops = re.findall(r"\b(ce|mn|mx)\((\d+), (\d+)\);", body)
    return [(k, int(i), int(j)) for k, i, j in ops]


def _run(ops, v):
    v = list(v)
    for k, i, j in ops:
        a, b = v[i], v[j]
        if k in ("ce", "mn"):
            v[i] = np.minimum(a, b)
        if k in ("ce", "mx"):
            v[j] = np.maximum(a, b)
    return v[12]


def test_network_shape():
    ops = _network()
    assert len(ops) == 113
    assert all(0 <= i < j < 25 for _, i, j in ops)


def test_zero_one_principle_exhaustive():
    ops = _network()
    n = 25
    nb = 1 << n
    idx = np.arange(nb, dtype=np.uint32)
    planes = [np.packbits(((idx >> k) & 1).astype(np.uint8), bitorder="little") for k in range(n)]
    got = _run_bits(ops, planes)
    pc = np.zeros(nb, dtype=np.uint8)
    for k in range(n):
        pc += ((idx >> k) & 1).astype(np.uint8)
    want = np.packbits((pc >= 13).astype(np.uint8), bitorder="little")
    assert np.array_equal(got, want)


def _run_bits(ops, planes):
    v = list(planes)
    for k, i, j in ops:
        a, b = v[i], v[j]
        if k in ("ce", "mn"):
            v[i] = a & b
        if k in ("ce", "mx"):
            v[j] = a | b
    return v[12]


def test_float_windows_with_ties():
    ops = _network()
    rng = np.random.default_rng(7)
    m = 20000
    for levels in (3, 7, 1000, None):
        x = rng.normal(size=(25, m)).astype(np.float32)
        if levels is not None:
            x = np.round(x * levels / 4) / (levels / 4)
        got = _run(ops, [x[k] for k in range(25)])
        want = np.sort(x, axis=0)[12]
        np.testing.assert_array_equal(got, want)

