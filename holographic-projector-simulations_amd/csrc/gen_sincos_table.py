"""Writes holo_sincos_table.inc, the table of hm_sincos (holo_math.hpp):
(sin, cos)(j pi / 128) for j < 256, each rounded to the nearest double from
200-bit mpmath values (sinpi/cospi, so the zeros are exact).  Also prints the
truncation bounds of the Taylor polynomials hm_sincos uses on |r| <= pi/256
and checks the table's odd/even symmetry that the FP64 kernel relies on.

    python gen_sincos_table.py
"""
import os

import mpmath as mp

mp.mp.prec = 200
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = []
    S, C = [], []
    for j in range(256):
        x = mp.mpf(j) / 128
        s, c = float(mp.sinpi(x)), float(mp.cospi(x))  # mpf -> double rounds to nearest
        S.append(s)
        C.append(c)
        rows.append(f"    {s.hex()}, {c.hex()},")
    for j in range(1, 256):  # sin(-x) = -sin x, cos(-x) = cos x, bit for bit
        assert S[256 - j] == -S[j] and C[256 - j] == C[j]
    with open(os.path.join(HERE, "holo_sincos_table.inc"), "w") as f:
        f.write("// (sin, cos)(j pi / 128), j < 256, correctly rounded; written by gen_sincos_table.py\n")
        f.write("\n".join(rows) + "\n")
    r = mp.pi / 256
    print("sin r - (r - r^3/6 + r^5/120) <=", mp.nstr(r ** 7 / 5040, 3))
    print("cos r - (1 - r^2/2 + r^4/24 - r^6/720) <=", mp.nstr(r ** 8 / 40320, 3))


if __name__ == "__main__":
    main()
