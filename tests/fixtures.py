"""Readers for the reference's Octave-ASCII fixture files (tests/golden/).

The format is the one read by read_{vector,matrix}_from_txt_file in the
reference (src/matlab_utilities.hpp:125-137): '#' header lines naming the
type and shape, then whitespace-separated values row by row.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_octave(path):
    rows = cols = None
    values = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("#"):
                if line.startswith("# rows:"):
                    rows = int(line.split(":")[1])
                elif line.startswith("# columns:"):
                    cols = int(line.split(":")[1])
                continue
            values.extend(float(v) for v in line.split())
    a = np.array(values)
    if rows is not None and cols is not None:
        a = a.reshape(rows, cols)
    return a


def golden(*parts):
    return os.path.join(GOLDEN, *parts)
