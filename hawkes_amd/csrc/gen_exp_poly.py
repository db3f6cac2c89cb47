"""Generates hph_exp_coeffs.h: constants of the FP64 exp and exp2 used in the pair kernels.

exp (Cody & Waite argument reduction + polynomial, W. J. Cody and W. Waite,
"Software Manual for the Elementary Functions", 1980):
    x = k*ln2 + r,   k = nearest integer to x/ln2,   |r| <= ln2/2 (+ rounding slack)
    exp(x) = 2^k * p(r),   p ~ e^r on [-A, A]
ln2 is split as LN2_HI + LN2_LO with LN2_HI holding 32 significant bits, so k*LN2_HI is exact
for |k| < 2^21.
exp2 (for exponents already scaled by log2 e):
    y = k + f,   k = nearest integer to y,   f = y - k exact,   2^y = 2^k * q(f),   q ~ 2^f
on [-1/2, 1/2].
exp2 of a non-negative exponent (background terms of the base-2 pair form, 2^-y with y >= 0):
    y = i + f,   i = trunc(y),   f = fract(y) in [0, 1) exact,   2^-y = 2^-i * s(f),   s ~ 2^-f
on [0, 1] (the same relative error as q: 2^-f on [0, 1] is 2^-1/2 times 2^g on [-1/2, 1/2]).
exp2 of a non-negative exponent from a table of S = 2^TAB_BITS entries (the FP64 background terms
of the base-2 pair form, whose exponent arrives scaled by S: yS = S y >= 0):
    -trunc(yS) = S a + b (b in [0, S - 1]),   f = fract(yS) in [0, 1) exact,
    2^-y = 2^a * M[b] * u(f),   M[b] = 2^(b/S) correctly rounded,   u ~ 2^(-f/S) on [0, 1].
With S = 256, u has degree 4 (the interval is 1/256 of the one above); S = 32 needs degree 5.
All four polynomials have value 1 at 0 exactly and the other coefficients minimise
the weighted (relative) least-squares error on 3000 Chebyshev nodes at 40 digits, which lands
close to the minimax polynomial. The script then measures the max relative error of the
double-precision Horner evaluation with correctly rounded fma at 20000 points.

Usage: python hawkes_amd/csrc/gen_exp_poly.py [degree [table bits [table degree]]]
(defaults 10, 8, 4; writes hph_exp_coeffs.h beside itself)
"""
import os
import random
import sys

import mpmath as mp

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def fit(degree, A, scale=1, lo=None):
    """p(r) = 1 + sum c_k r^k ~ e^(scale r) on [-A, A] (on [lo, A] when lo is given)."""
    lo = -A if lo is None else lo
    mid, half = (A + lo) / 2, (A - lo) / 2
    nodes = [mid + half * mp.cos(mp.pi * (j + mp.mpf(0.5)) / 3000) for j in range(3000)]
    # unknowns c1..c_degree; minimise sum ((p(r) - e^(scale r)) / e^(scale r))^2
    rows, rhs = [], []
    for r in nodes:
        w = mp.exp(-scale * r)
        rows.append([w * r ** k for k in range(1, degree + 1)])
        rhs.append(w * (mp.exp(scale * r) - 1))
    M = mp.matrix(rows)
    b = mp.matrix(rhs)
    sol = mp.qr_solve(M, b)[0]
    return [mp.mpf(1)] + [sol[k] for k in range(degree)]


def fma_d(a, b, c):
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def horner(coef_d, r):
    p = coef_d[-1]
    for c in reversed(coef_d[:-1]):
        p = fma_d(p, r, c)
    return p


def main():
    degree = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    tbits = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    tdeg = int(sys.argv[3]) if len(sys.argv) > 3 else 4
    S = 1 << tbits
    ln2 = mp.log(2)
    ln2_hi = float(mp.nint(ln2 * 2 ** 32) / 2 ** 32)
    ln2_lo = float(ln2 - mp.mpf(ln2_hi))
    log2e = float(1 / ln2)
    A = ln2 / 2 * (1 + mp.mpf(2) ** -20)
    coef = fit(degree, A)
    coef_d = [float(c) for c in coef]
    random.seed(1)
    worst = 0.0
    for _ in range(20000):
        x = random.uniform(-740.0, 700.0)
        k = round(x * log2e)
        r = fma_d(-k, ln2_lo, fma_d(-k, ln2_hi, x))
        y = mp.ldexp(mp.mpf(horner(coef_d, r)), k)
        worst = max(worst, abs(float((y - mp.exp(x)) / mp.exp(x))))
    coef2 = fit(degree, mp.mpf(0.5), ln2)
    coef2_d = [float(c) for c in coef2]
    worst2 = 0.0
    for _ in range(20000):
        y = random.uniform(-1070.0, 1020.0)
        k = round(y)
        v = mp.ldexp(mp.mpf(horner(coef2_d, y - k)), k)
        worst2 = max(worst2, abs(float((v - mp.power(2, y)) / mp.power(2, y))))
    coef3 = fit(degree, mp.mpf(1), -ln2, mp.mpf(0))
    coef3_d = [float(c) for c in coef3]
    worst3 = 0.0
    for _ in range(20000):
        y = random.choice((random.uniform(0.0, 4.0), random.uniform(0.0, 1070.0)))
        i = int(y)
        v = mp.ldexp(mp.mpf(horner(coef3_d, y - i)), -i)
        worst3 = max(worst3, abs(float((v - mp.power(2, -y)) / mp.power(2, -y))))
    coef4 = fit(tdeg, mp.mpf(1), -ln2 / S, mp.mpf(0))
    coef4_d = [float(c) for c in coef4]
    mant = [float(mp.power(2, mp.mpf(b) / S)) for b in range(S)]
    worst4 = 0.0
    for _ in range(20000):
        y = random.choice((random.uniform(0.0, 4.0), random.uniform(0.0, 1020.0)))
        yS = float(mp.mpf(y) * S)
        ni = -int(yS)
        f = yS + ni
        t = mant[ni & (S - 1)]
        q = coef4_d[-1]
        for c in reversed(coef4_d[1:-1]):
            q = fma_d(q, f, c)
        v = mp.ldexp(mp.mpf(fma_d(t, float(mp.mpf(f) * mp.mpf(q)), t)), ni >> tbits)
        ref = mp.power(2, -mp.mpf(yS) / S)
        worst4 = max(worst4, abs(float((v - ref) / ref)))
    path = os.path.join(HERE, "hph_exp_coeffs.h")
    with open(path, "w") as f:
        f.write("// Generated by gen_exp_poly.py (degree %d); do not edit.\n" % degree)
        f.write("// Measured max relative error of the FP64 evaluation: exp %.3e, exp2 %.3e, exp2 of -y %.3e,\n"
                "// exp2 of -y from the table %.3e\n" % (worst, worst2, worst3, worst4))
        f.write("#pragma once\n")
        f.write("#define HPH_EXP_DEGREE %d\n" % degree)
        f.write("#define HPH_LOG2E %r\n" % log2e)
        f.write("#define HPH_LN2_HI %r\n" % ln2_hi)
        f.write("#define HPH_LN2_LO %r\n" % ln2_lo)
        for k, c in enumerate(coef_d):
            f.write("#define HPH_EXP_C%d %r\n" % (k, c))
        for k, c in enumerate(coef2_d):
            f.write("#define HPH_EXP2_C%d %r\n" % (k, c))
        for k, c in enumerate(coef3_d):
            f.write("#define HPH_EXP2N_C%d %r\n" % (k, c))
        f.write("#define HPH_EXP2T_BITS %d\n" % tbits)
        f.write("#define HPH_EXP2T_DEGREE %d\n" % tdeg)
        for k, c in enumerate(coef4_d):
            f.write("#define HPH_EXP2T_C%d %r\n" % (k, c))
        # u(f) = 1 + f q(f): the coefficients of q, lowest first
        f.write("#define HPH_EXP2T_Q %s\n" % ", ".join("HPH_EXP2T_C%d" % k for k in range(1, tdeg + 1)))
        for j, m in enumerate(mant):
            f.write("#define HPH_EXP2T_M%d %r\n" % (j, m))
        f.write("#define HPH_EXP2T_TABLE %s\n" % ", ".join("HPH_EXP2T_M%d" % j for j in range(S)))
    print("degree", degree, "table", S, "x degree", tdeg, "max rel err exp %.3e exp2 %.3e exp2 of -y %.3e table %.3e" % (worst, worst2, worst3, worst4),
          "->", path)


if __name__ == "__main__":
    main()
