"""Writes tests/golden/gaunt_factor_vanhoof.npy from the reference's Gaunt
table (gaunt_factor.dat, read by gaunt_factor.cpp:30-41 through
data_table_loader.hpp:39-97): '#' comment lines, then 146 rows (ln u from -16
in steps of 0.2) of 81 values (ln gamma^2 from -6 in steps of 0.2).  The values
are stored as float32, as the reference's loader reads them
(data_table_loader<device_scalar>), row-major.

    python tests/golden/make_gaunt_fixture.py [path/to/gaunt_factor.dat]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/gaunt_factor.dat"
OUT = os.path.join(HERE, "gaunt_factor_vanhoof.npy")


def main():
    rows = []
    with open(SRC) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                rows.append([float(v) for v in line.split()])
    tab = np.asarray(rows, dtype=np.float32)
    assert tab.shape == (146, 81), tab.shape
    np.save(OUT, tab, allow_pickle=False)
    print(f"{OUT}: {tab.shape} float32, min {tab.min():.6g} max {tab.max():.6g}")


if __name__ == "__main__":
    main()
