"""Derives the FP64 sin/cos polynomial coefficients used in holo_math.hpp.

sin r = r + r^3 * P(r^2),  cos r = 1 + r^2 * Q(r^2)  on |r| <= pi/2 (+ margin),
with P of degree 6 and Q of degree 6 in u = r^2, fitted by weighted least
squares at 50-digit precision on Chebyshev nodes in u, weights chosen so
that the fitted quantity is the absolute error of sin r and cos r.  Prints
the coefficients as hex doubles and the maximum absolute error of the
double-rounded coefficients, evaluated in mpmath on a dense grid.

    python gen_sincos_coeffs.py
"""
import mpmath as mp

mp.mp.dps = 50
HALF_PI = mp.pi / 2
U_MAX = (HALF_PI * (1 + mp.mpf("1e-6"))) ** 2


def fit(target, weight, degree, nodes=200):
    """Least-squares coefficients c_0..c_degree of sum c_k u^k ~ target(u)."""
    us = [U_MAX * (1 - mp.cos(mp.pi * (j + mp.mpf(0.5)) / nodes)) / 2 for j in range(nodes)]
    A = mp.matrix(nodes, degree + 1)
    b = mp.matrix(nodes, 1)
    for i, u in enumerate(us):
        w = weight(u)
        for k in range(degree + 1):
            A[i, k] = w * u ** k
        b[i] = w * target(u)
    x = mp.lu_solve(A.T * A, A.T * b)
    return [x[k] for k in range(degree + 1)]


def sin_target(u):  # (sin r - r) / r^3
    r = mp.sqrt(u)
    return (mp.sin(r) - r) / r ** 3 if u > 0 else mp.mpf(-1) / 6


def cos_target(u):  # (cos r - 1) / r^2
    r = mp.sqrt(u)
    return (mp.cos(r) - 1) / u if u > 0 else mp.mpf(-1) / 2


def main():
    P = fit(sin_target, lambda u: mp.sqrt(u) ** 3, 6)
    Q = fit(cos_target, lambda u: u, 6)
    Pd = [float(c) for c in P]
    Qd = [float(c) for c in Q]
    err_s = err_c = mp.mpf(0)
    for j in range(4001):
        r = HALF_PI * (1 + mp.mpf("1e-6")) * j / 4000
        u = r * r
        ps = sum(mp.mpf(c) * u ** k for k, c in enumerate(Pd))
        pc = sum(mp.mpf(c) * u ** k for k, c in enumerate(Qd))
        err_s = max(err_s, abs(r + r ** 3 * ps - mp.sin(r)))
        err_c = max(err_c, abs(1 + u * pc - mp.cos(r)))
    print("sin P (u^0 .. u^6):", ", ".join(c.hex() for c in Pd))
    print("cos Q (u^0 .. u^6):", ", ".join(c.hex() for c in Qd))
    print(f"max abs error with double coefficients: sin {float(err_s):.3e}, cos {float(err_c):.3e}")


if __name__ == "__main__":
    main()
