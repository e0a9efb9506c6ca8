"""Write tests/golden/colormap_table.npz: the reference's 256 x 3 colour table
(Python/colormap.py:27-284, ``colormap_raw``, R, G, B) as data, read from the
source text (the module is not imported: it needs np.int, gone from NumPy 2).

Run from the repository root, where the reference is checked out at
/root/reference: python tests/golden/make_colormap_fixture.py
"""

import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/Python/colormap.py"


def main():
    text = open(SRC).read()
    start = text.index("colormap_raw = np.array([")
    end = text.index("]])", start)
    rows = re.findall(r"\[([^\[\]]+)\]", text[start:end + 2])
    table = np.array([[float(v) for v in r.split(",")] for r in rows])
    assert table.shape == (256, 3), table.shape
    np.savez_compressed(os.path.join(HERE, "colormap_table.npz"), colormap_raw=table)
    print("wrote", table.shape)


if __name__ == "__main__":
    main()
