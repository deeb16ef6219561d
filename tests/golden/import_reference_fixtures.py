#!/usr/bin/env python3
"""Convert the reference's regression fixtures into tests/golden data.

The reference's tests/test_landau.py compares `last_0.dat` after 100 steps of
tests/data/simple_landau.toml with tests/data/reference/landau_{euler,
pseudospectral}_{1D,2D,3D}.dat.  This script parses those six text files
(values written at the default 6-significant-digit ostream precision) into
tests/golden/reference_landau_fixtures.npz (one float64 array per file plus
its header line), and copies the input TOML next to it.  Formatting a parsed
value with '%g' reproduces the reference's token exactly, so tests compare
'%g' strings.  Run in the development container (reads /root/reference).
"""
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/data"


def main():
    out = {}
    for name in ("euler", "pseudospectral"):
        for d in (1, 2, 3):
            with open(os.path.join(SRC, "reference", f"landau_{name}_{d}D.dat")) as f:
                header, body = f.read().split("\n", 1)
            out[f"{name}_{d}D"] = np.array(body.split(), dtype=np.float64)
            out[f"{name}_{d}D_header"] = np.array(header)
    np.savez_compressed(os.path.join(HERE, "reference_landau_fixtures.npz"), **out)
    shutil.copy(os.path.join(SRC, "simple_landau.toml"), os.path.join(HERE, "simple_landau.toml"))


if __name__ == "__main__":
    main()
