"""Writes gl25.h: the 25-point Gauss-Legendre rule on [-1, 1] (nodes ascending, weights),
each value the double nearest to the exact one (mpmath at 60 digits).  The reference takes
the same rule from GSL's gsl_integration_fixed_legendre (S/Utilities.C:112-122).
Run: python amcafe_amd/csrc/gen_gl25.py"""
import os

import mpmath as mp

mp.mp.dps = 60
N = 25


def legendre_rule(n):
    xs, ws = [], []
    for i in range(1, n + 1):
        x = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))  # initial guess, descending
        for _ in range(100):
            p0, p1 = mp.mpf(1), x
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = n * (x * p1 - p0) / (x * x - 1)
            dx = p1 / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-55):
                break
        p0, p1 = mp.mpf(1), x
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = n * (x * p1 - p0) / (x * x - 1)
        if abs(x) < mp.mpf(10) ** (-50):
            x = mp.mpf(0)  # the middle root of the odd P25 is exactly 0
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    order = sorted(range(n), key=lambda i: xs[i])
    return [xs[i] for i in order], [ws[i] for i in order]


def main():
    xs, ws = legendre_rule(N)
    h = lambda v: float(v).hex()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gl25.h")
    with open(out, "w") as f:
        f.write("// gl25.h -- written by gen_gl25.py: 25-point Gauss-Legendre rule on [-1, 1],\n")
        f.write("// nodes ascending, each value the double nearest to the exact one.\n#pragma once\n")
        f.write("// Initializer lists, for host and device arrays alike.\n")
        f.write("#define AMC_GL25_N 25\n")
        f.write("#define AMC_GL25_X { \\\n" + ", \\\n".join("    " + h(x) for x in xs) + "}\n")
        f.write("#define AMC_GL25_W { \\\n" + ", \\\n".join("    " + h(w) for w in ws) + "}\n")


if __name__ == "__main__":
    main()
