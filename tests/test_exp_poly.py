"""Accuracy of the FP64 exponentials of the pair kernels (hawkes_amd/csrc/hph_device.h), pinned
on the CPU: the constants of hph_exp_coeffs.h evaluated as the kernels evaluate them (the same
reductions, Horner with correctly rounded fma, emulated at 60 digits) against mpmath at 60
digits. exp_fp64 (Cody-Waite, natural form), exp2_fp64 (shifter, exponents of either sign) and
exp2_neg_fp64 (2^-y for y >= 0 from trunc and fract: the background terms of the base-2 form).
A dropped or mistyped coefficient, a wrong sign of the 2^-f polynomial or an off-by-one in the
reduction moves the error far above the bound.
"""
import os
import re

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "hawkes_amd", "csrc", "hph_exp_coeffs.h")
mp.mp.dps = 60


def constants():
    c = {}
    with open(HEADER) as f:
        for line in f:
            m = re.match(r"#define (HPH_\w+) (\S+)", line)
            if m:
                c[m.group(1)] = float(m.group(2))
    return c


def fma(a, b, c):
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def horner(coef, x):
    p = coef[-1]
    for cc in reversed(coef[:-1]):
        p = fma(p, x, cc)
    return p


def poly(c, prefix):
    deg = int(c["HPH_EXP_DEGREE"])
    return [c["%s%d" % (prefix, k)] for k in range(deg + 1)]


def ldexp(p, k):
    """v_ldexp_f64: p 2^k rounded to FP64 (subnormal results rounded, underflow to +0)."""
    return float(mp.ldexp(mp.mpf(p), k))


def exp_fp64(c, x):
    x = max(x, -1000.0)
    shifter = 6755399441055744.0
    kf = fma(x, c["HPH_LOG2E"], shifter)
    k = kf - shifter
    r = fma(-k, c["HPH_LN2_HI"], x)
    r = fma(-k, c["HPH_LN2_LO"], r)
    return ldexp(horner(poly(c, "HPH_EXP_C"), r), int(k))


def exp2_fp64(c, y):
    shifter = 6755399441055744.0
    kf = y + shifter
    k = kf - shifter
    f = y - k
    return ldexp(horner(poly(c, "HPH_EXP2_C"), f), int(k))


def exp2_neg_fp64(c, y):
    i = int(y)           # v_cvt_i32_f64 of -y gives -trunc(y)
    f = y - np.floor(y)  # v_fract_f64, exact for y >= 0
    return ldexp(horner(poly(c, "HPH_EXP2N_C"), f), -i)


def rel(v, ref):
    return abs(float((mp.mpf(v) - ref) / ref))


def test_exp_fp64():
    c = constants()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-700.0, 700.0, 300), rng.uniform(-3.0, 3.0, 300), [0.0, -1e-300, 1e-20]])
    worst = max(rel(exp_fp64(c, float(x)), mp.exp(mp.mpf(float(x)))) for x in xs)
    assert worst < 6e-16, worst
    assert exp_fp64(c, 0.0) == 1.0
    assert exp_fp64(c, -2000.0) == 0.0  # clamped to -1000: underflows to +0 as IEEE exp does


def test_exp2_fp64():
    c = constants()
    rng = np.random.default_rng(2)
    ys = np.concatenate([rng.uniform(-1020.0, 1020.0, 300), rng.uniform(-2.0, 2.0, 300), [0.0, 0.5, -0.5]])
    worst = max(rel(exp2_fp64(c, float(y)), mp.power(2, mp.mpf(float(y)))) for y in ys)
    assert worst < 6e-16, worst


def test_exp2_neg_fp64():
    c = constants()
    rng = np.random.default_rng(3)
    ys = np.concatenate([rng.uniform(0.0, 1020.0, 300), rng.uniform(0.0, 4.0, 300),
                         [0.0, 1.0, 2.0, 1.0 - 2.0 ** -53, 0.5, 1e-300]])
    worst = max(rel(exp2_neg_fp64(c, float(y)), mp.power(2, -mp.mpf(float(y)))) for y in ys)
    assert worst < 7e-16, worst
    assert exp2_neg_fp64(c, 0.0) == 1.0  # the self pair's term is exactly 1
    assert exp2_neg_fp64(c, 3.0) == 0.125
    assert exp2_neg_fp64(c, 1100.0) == 0.0
