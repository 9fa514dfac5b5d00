"""Writes tests/golden/kinetic_slice.npz from the reference's kinetic
fluorescence data (tensorly/datasets/data/Kinetic.npy and
Kinetic_missing.npy, loaded by tensorly/datasets/data_imports.py:879-905;
public-use data of Nikolajsen, Booksh, Hansen and Bro, 2003).

The fixture keeps every second index of modes 0 and 3 (32 x 12 x 10 x 30 of
64 x 12 x 10 x 60).  Every value of the data is an integer multiple of 1/3
(|3x| <= 8318), so ``x3 = 3 x`` is stored exactly as int16; missing entries
are the zeros the reference's loader leaves, and their positions are stored
as ``missing``.  Run in a checkout where the reference sits at
/root/reference; the data files are read with numpy.load(allow_pickle=False).
"""
import os

import numpy as np

SRC = "/root/reference/tensorly/datasets/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kinetic_slice.npz")


def main():
    x = np.load(os.path.join(SRC, "Kinetic.npy"), allow_pickle=False)
    missing = np.load(os.path.join(SRC, "Kinetic_missing.npy"), allow_pickle=False)
    s = x[::2, :, :, ::2]
    x3 = np.round(s * 3.0)
    assert np.array_equal(x3 / 3.0, s), "values are not multiples of 1/3"
    assert np.abs(x3).max() < 2 ** 15
    np.savez_compressed(OUT, x3=x3.astype(np.int16), missing=missing[::2, :, :, ::2])
    print(f"wrote {OUT}: shape {s.shape}")


if __name__ == "__main__":
    main()
