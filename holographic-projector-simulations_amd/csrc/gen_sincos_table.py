"""Writes holo_sincos_table.inc, the table of hm_sincos (holo_math.hpp):
(sin, cos)(j pi / 1024) for j < 2048, each rounded to the nearest double from
200-bit mpmath values (sinpi/cospi, so the zeros are exact).  Also prints the
truncation bounds of the Taylor polynomials hm_sincos uses on |r| <= pi/2048
and checks the table's odd/even symmetry that the FP64 kernel relies on.

    python gen_sincos_table.py
"""
import os

import mpmath as mp

mp.mp.prec = 200
STEPS = 2048  # table entries per turn
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = []
    S, C = [], []
    for j in range(STEPS):
        x = mp.mpf(2 * j) / STEPS
        s, c = float(mp.sinpi(x)), float(mp.cospi(x))  # mpf -> double rounds to nearest
        S.append(s)
        C.append(c)
        rows.append(f"    {s.hex()}, {c.hex()},")
    for j in range(1, STEPS):  # sin(-x) = -sin x, cos(-x) = cos x, bit for bit
        assert S[STEPS - j] == -S[j] and C[STEPS - j] == C[j]
    with open(os.path.join(HERE, "holo_sincos_table.inc"), "w") as f:
        f.write(f"// (sin, cos)(j pi / {STEPS // 2}), j < {STEPS}, correctly rounded; written by gen_sincos_table.py\n")
        f.write("\n".join(rows) + "\n")
    r = mp.pi / STEPS
    print("sin r - (r - r^3/6) <=", mp.nstr(r ** 5 / 120, 3))
    print("cos r - (1 - r^2/2 + r^4/24) <=", mp.nstr(r ** 6 / 720, 3))


if __name__ == "__main__":
    main()
