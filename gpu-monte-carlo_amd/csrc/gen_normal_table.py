"""Prints the cubic pieces of the normal quantile table in mc_fastmath.h
(kNormalQ; the oracle's OFM_NORMAL_Q is the same table).

Q(w) = -Phi^-1(w) is sampled at w = (2k + 1) 2^-23, k < 2^21.  Cell
8 E + i (E = 0..21, i = 0..7) covers 2k + 1 in [2^E (1 + i/8), 2^E (1 + (i+1)/8))
with t in [0, 1) across the cell; its four float32 coefficients are the float64
least-squares cubic in t through Q at 64 Chebyshev points of the cell, rounded
to float32.  Usage: python gen_normal_table.py > rows.txt
"""
import numpy as np
from scipy.special import ndtri


def cells():
    t = 0.5 - 0.5 * np.cos(np.pi * (np.arange(64) + 0.5) / 64)
    for c in range(22 * 8):
        e, i = divmod(c, 8)
        w = 2.0**e * (1 + (i + t) / 8) * 2.0**-23
        yield np.polynomial.polynomial.polyfit(t, -ndtri(w), 3).astype(np.float32)


def hexf(x):
    m, e = float(x).hex().split("p")
    if "." in m:
        m = m.rstrip("0").rstrip(".")
    return f"{m}p{e}f"


if __name__ == "__main__":
    for row in cells():
        print("    {" + ", ".join(hexf(x) for x in row) + "},")
