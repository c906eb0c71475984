"""Writes twiddles.hpp: cos/sin(2*pi*m/120), m = 0..119, rounded to the nearest double.

120 = lcm(3, 4, 5, 6, 8, 10, 12) covers every DFT length used by the 6D grid kernels
(nGrid = 2N for N = 1..6).  The values are computed with 50-digit decimal arithmetic
from the Taylor series on [0, pi/4] and unfolded by exact symmetries, so every entry is
the correctly rounded double of the true value.
"""
from decimal import Decimal, getcontext
import os

getcontext().prec = 60


def _pi():
    # Machin: pi = 16 atan(1/5) - 4 atan(1/239)
    def atan_inv(x):
        x = Decimal(x)
        s, t, n, k = Decimal(0), 1 / x, 1, 0
        while True:
            term = t / n
            if term == 0 or abs(term) < Decimal(10) ** -58:
                break
            s += term if k % 2 == 0 else -term
            t /= x * x
            n += 2
            k += 1
        return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)


PI = _pi()


def _sincos_small(a):
    s, c = Decimal(0), Decimal(0)
    term = Decimal(1)
    for n in range(0, 80):
        if n > 0:
            term = term * a / n
        r = n % 4
        if r == 0:
            c += term
        elif r == 1:
            s += term
        elif r == 2:
            c -= term
        else:
            s -= term
    return s, c


def sincos_m120(m):
    """Exact-symmetry evaluation of sin/cos(2*pi*m/120)."""
    m %= 120
    # reduce to the first octant m in [0, 15]
    sign_s, sign_c, swap = 1, 1, False
    if m >= 60:
        m -= 60
        sign_s, sign_c = -sign_s, -sign_c
    if m > 30:
        m = 60 - m
        sign_c = -sign_c
    if m > 15:
        m = 30 - m
        swap = True
    s, c = _sincos_small(2 * PI * m / 120)
    if swap:
        s, c = c, s
    return sign_s * s, sign_c * c


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twiddles.hpp")
    cs, ss = [], []
    for m in range(120):
        s, c = sincos_m120(m)
        cs.append(repr(float(c)) if float(c) != 0 else "0.0")
        ss.append(repr(float(s)) if float(s) != 0 else "0.0")
    with open(out, "w") as f:
        f.write("// Generated by gen_twiddles.py: cos/sin(2*pi*m/120) rounded to nearest double.\n")
        f.write("#pragma once\n\nnamespace fame6d {\n\n")
        f.write("constexpr double kCos120[120] = {\n  " + ",\n  ".join(cs) + "};\n\n")
        f.write("constexpr double kSin120[120] = {\n  " + ",\n  ".join(ss) + "};\n\n")
        f.write("}  // namespace fame6d\n")


if __name__ == "__main__":
    main()
