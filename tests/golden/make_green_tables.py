"""Extracts the Mathematica value tables of the reference's Green/Bessel test
(src/test/test_green_bessel_data.h) into tests/golden/green_tables.npz.

Each table row is [order-or-dimension, Re(arg), Im(arg), Re(value), Im(value)];
the arrays are stored as float64 exactly as written in the header.  Run in a
container that holds the reference tree:  python tests/golden/make_green_tables.py
"""
import os
import re
import sys

import numpy as np

SRC = os.environ.get("MSMODEL_REF", "/root/reference") + "/src/test/test_green_bessel_data.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "green_tables.npz")
TABLES = ["BESSEL_K_DATA", "GREEN_DATA", "GREEN_IMAG_DATA", "DK_GREEN_DATA"]


def main():
    text = open(SRC).read()
    out = {}
    for name in TABLES:
        m = re.search(r"double\s+%s\s*\[\]\s*\[5\]\s*=\s*\{(.*?)\};" % name, text, re.S)
        if not m:
            sys.exit("table %s not found" % name)
        nums = re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", m.group(1))
        arr = np.array([float(x) for x in nums], dtype=np.float64).reshape(-1, 5)
        out[name.lower()] = arr
        print(name, arr.shape)
    np.savez_compressed(OUT, **out)


if __name__ == "__main__":
    main()
