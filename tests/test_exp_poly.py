"""Accuracy of the FP64 exponentials of the pair kernels (hawkes_amd/csrc/hph_device.h), pinned
on the CPU: the constants of hph_exp_coeffs.h evaluated as the kernels evaluate them (the same
reductions, Horner with correctly rounded fma, emulated at 60 digits) against mpmath at 60
digits. exp_fp64 (Cody-Waite, natural form), exp2_fp64 (shifter, exponents of either sign) and
exp2_neg_fp64 (2^-y for y >= 0 from trunc and fract) and exp2_neg_tab (2^-y from 256 y: a
256-entry table of 2^(b/256), a degree-4 polynomial and v_ldexp_f64; the FP64 background terms of
the base-2 form), and exp2_neg_ktab (the same with the power of two added into the table entry's
exponent field) bitwise against exp2_neg_tab over its range.
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
            m = re.match(r"#define (HPH_\w+) ([-+0-9.eE]+)$", line.strip())
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


def tab_bits(c):
    return int(c["HPH_EXP2T_BITS"])


def exp2_table(c):
    """fill_exp2_table: entry b = 2^(b/S), S = 2^HPH_EXP2T_BITS."""
    return [c["HPH_EXP2T_M%d" % b] for b in range(1 << tab_bits(c))]


def exp2t_fq(c, f):
    """f q(f), q by Horner from its top coefficient (hph_device.h exp2t_fq)."""
    deg = int(c["HPH_EXP2T_DEGREE"])
    q = c["HPH_EXP2T_C%d" % deg]
    for k in range(deg - 1, 0, -1):
        q = fma(q, f, c["HPH_EXP2T_C%d" % k])
    return float(mp.mpf(f) * mp.mpf(q))


def exp2_neg_tab(c, tab, yS):
    """2^(-yS/S): -trunc(yS) = S a + b, f = fract(yS), ldexp(tab[b] + tab[b] (f q(f)), a)."""
    bits = tab_bits(c)
    ni = -int(yS)  # v_cvt_i32_f64 of -yS
    f = yS - np.floor(yS)
    t = tab[ni & ((1 << bits) - 1)]  # Python's & and >> on negative ints: two's complement, arithmetic shift
    return ldexp(fma(t, exp2t_fq(c, f), t), ni >> bits)


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


def test_exp2_table_mantissas():
    """The 256 table entries are 2^(b/256) correctly rounded, and hph_device.h's fixed choices
    (the byte index, the scale constant) match the header's table size."""
    c = constants()
    assert tab_bits(c) == 8
    S = 1 << tab_bits(c)
    for b in range(S):
        assert c["HPH_EXP2T_M%d" % b] == float(mp.power(2, mp.mpf(b) / S)), b


def test_exp2_neg_tab():
    c = constants()
    tab = exp2_table(c)
    S = float(len(tab))
    rng = np.random.default_rng(4)
    ys = np.concatenate([rng.uniform(0.0, 1020.0, 300), rng.uniform(0.0, 4.0, 300),
                         [0.0, 1.0 / S, 1.0, 2.0, 95.0, 1e-300, 1.0 - 2.0 ** -53, (S - 1) / S, (S + 1) / S,
                          1.0 / 32, 31.0 / 32]])
    worst = 0.0
    for y in ys:
        yS = float(mp.mpf(float(y)) * mp.mpf(S))
        worst = max(worst, rel(exp2_neg_tab(c, tab, yS), mp.power(2, -mp.mpf(yS) / mp.mpf(S))))
    assert worst < 3e-16, worst
    assert exp2_neg_tab(c, tab, 0.0) == 1.0  # the self pair's term is exactly 1
    assert exp2_neg_tab(c, tab, 3.0 * S) == 0.125  # 2^-3
    assert exp2_neg_tab(c, tab, S * 1100) == 0.0  # below the FP64 range, as IEEE exp
    assert exp2_neg_tab(c, tab, 2.0 ** 31 - 1) == 0.0


def exp2_ktable(c, tab):
    """fill_exp2_table with ktab: the bit pattern of entry b with b << (20 - bits) taken off its
    high word."""
    sh = 32 + 20 - tab_bits(c)
    return [int(np.array([t]).view(np.uint64)[0]) - (b << sh) for b, t in enumerate(tab)]


def exp2_neg_ktab(c, ktab, yS):
    """exp2_neg_ktab: as exp2_neg_tab, with (-trunc(yS)) << (20 - bits) added to the table entry's
    high word (32-bit wrap-around) in place of the final ldexp."""
    tb = tab_bits(c)
    ni = -int(yS)
    f = yS - np.floor(yS)
    bits = ktab[ni & ((1 << tb) - 1)]
    hi = ((bits >> 32) + ((ni << (20 - tb)) & 0xFFFFFFFF)) & 0xFFFFFFFF
    t = float(np.array([(hi << 32) | (bits & 0xFFFFFFFF)], dtype=np.uint64).view(np.float64)[0])
    return fma(t, exp2t_fq(c, f), t)


def test_exp2_neg_ktab_is_bitwise_exp2_neg_tab():
    """Over its whole range (0 <= yS <= kExp2KTabMaxY) the exponent-in-table variant returns
    exp2_neg_tab's doubles bit for bit: every entry b, both ends of the range, integer and
    fractional y32. Its range covers every background exponent up to the host's bound (ktab_ok in
    hph_engine.cpp), and the smallest result is a normal double."""
    c = constants()
    tab = exp2_table(c)
    ktab = exp2_ktable(c, tab)
    S = len(tab)
    ymax = device_constants()["kExp2KTabMaxY"]
    assert ymax == S * 1011.0
    with open(os.path.join(HERE, "..", "hawkes_amd", "csrc", "hph_engine.cpp")) as f:
        src = f.read()
    body = src[src.index("static bool ktab_ok("):]
    bound = float(re.search(r"if \(!\(aB <= ([0-9.]+)\)\) return false;", body).group(1))
    assert bound * S * float(mp.log(mp.e, 2)) < ymax
    rng = np.random.default_rng(9)
    ys = np.concatenate([rng.uniform(0.0, ymax, 400), rng.uniform(0.0, 2.0 * S, 200),
                         np.arange(0.0, S + 70.0), [ymax, ymax - 1e-9, ymax - S + 0.5, 0.5, 1e-300, S - 1e-6]])
    for yS in ys:
        yS = float(yS)
        a, b = exp2_neg_ktab(c, ktab, yS), exp2_neg_tab(c, tab, yS)
        assert a == b and np.float64(a).tobytes() == np.float64(b).tobytes(), (yS, a, b)
    assert exp2_neg_ktab(c, ktab, ymax) >= 2.0 ** -1022


def exp2_neg_tab_on_ktab(c, ktab, yS):
    """exp2_neg_tab_on_ktab: the ktab entry with b << (20 - bits) added back to its high word, then
    exp2_neg_tab's ldexp form."""
    tb = tab_bits(c)
    ni = -int(yS)
    f = yS - np.floor(yS)
    b = ni & ((1 << tb) - 1)
    bits = ktab[b]
    hi = ((bits >> 32) + (b << (20 - tb))) & 0xFFFFFFFF
    t = float(np.array([(hi << 32) | (bits & 0xFFFFFFFF)], dtype=np.uint64).view(np.float64)[0])
    return ldexp(fma(t, exp2t_fq(c, f), t), ni >> tb)


def test_exp2_neg_tab_on_ktab_is_bitwise_exp2_neg_tab():
    """The excitation terms of the exponent-in-table kernels read exp2_neg_ktab's table but may
    lie far past its range: restoring each entry gives exp2_neg_tab's doubles bit for bit, past
    the range too, down to the subnormals and +0."""
    c = constants()
    tab = exp2_table(c)
    ktab = exp2_ktable(c, tab)
    S = float(len(tab))
    rng = np.random.default_rng(10)
    ys = np.concatenate([rng.uniform(0.0, S * 1100.0, 300), rng.uniform(0.0, 2.0 * S, 200),
                         np.arange(0.0, S + 8.0), [S * 1011.0, S * 1021.5, S * 1050.0, S * 1074.9, 2.0 ** 31 - 1]])
    for yS in ys:
        yS = float(yS)
        a, b = exp2_neg_tab_on_ktab(c, ktab, yS), exp2_neg_tab(c, tab, yS)
        assert np.float64(a).tobytes() == np.float64(b).tobytes(), (yS, a, b)


# ---- double-double constants of the base-2 form (hph_device.h) ----

DEVICE_H = os.path.join(HERE, "..", "hawkes_amd", "csrc", "hph_device.h")


def device_constants():
    """The `constexpr double kName = value` constants of hph_device.h."""
    c = {}
    with open(DEVICE_H) as f:
        text = f.read()
    for m in re.finditer(r"\b(k\w+)\s*=\s*([-+0-9.eE]+)\s*[,;]", text):
        c[m.group(1)] = float(m.group(2))
    return c


DD_VALUES = {
    "kSqrtHalfLog2e": lambda: mp.sqrt(mp.log(mp.e, 2) / 2),
    "kTwoLn2": lambda: 2 * mp.log(2),
    "kLog2e": lambda: 1 / mp.log(2),
    "kSqrt128Log2e": lambda: mp.sqrt(128 * mp.log(mp.e, 2)),
}


def test_double_double_constants():
    """Each (Hi, Lo) pair is the value rounded to a double plus the rounded remainder, so that
    Hi + Lo carries it to ~2^-106 relative; and 2 ln 2 is 1 / (sqrt(log2(e) / 2))^2 (what the
    finalize uses to undo the scaling of squared coordinates and times)."""
    c = device_constants()
    for name, value in DD_VALUES.items():
        v = value()
        hi, lo = c[name + "Hi"], c[name + "Lo"]
        assert hi == float(v), name
        assert lo == float(v - mp.mpf(hi)), name
        assert abs(mp.mpf(hi) + mp.mpf(lo) - v) <= mp.mpf(2) ** -104 * abs(v), name
    assert abs(DD_VALUES["kTwoLn2"]() * DD_VALUES["kSqrtHalfLog2e"]() ** 2 - 1) < mp.mpf(10) ** -50
    # the table form's scale: 256 times the squared scale, so 2 ln 2 / 256 undoes it
    assert abs(DD_VALUES["kTwoLn2"]() / 256 * DD_VALUES["kSqrt128Log2e"]() ** 2 - 1) < mp.mpf(10) ** -50


def time_scale_dd(c, tau, name="kSqrtHalfLog2e"):
    """hph_device.h time_scale_dd (time_scale_tab_dd with name kSqrt128Log2e): |tau| times the
    constant as a double-double (two-product)."""
    a = abs(tau)
    hi = float(mp.mpf(a) * mp.mpf(c[name + "Hi"]))
    err = fma(a, c[name + "Hi"], -hi)
    return hi, fma(a, c[name + "Lo"], err)


def test_scaling_round_trip():
    """A squared difference scaled by the double-double st and brought back by from_base2_sq
    (m 2 ln 2 / tau^2) returns the natural square to a few ulp: no systematic bias from a rounded
    scale (which is what a single rounded st would leave on every term of a dataset)."""
    c = device_constants()
    rng = np.random.default_rng(11)
    # (scale constant, unit of from_base2_sq): FP32-bulk's base-2 form and the FP64 table form
    for name, unit in (("kSqrtHalfLog2e", 1.0), ("kSqrt128Log2e", 1.0 / 256)):
        worst_scale = worst_back = 0.0
        for tau in np.concatenate([rng.uniform(1e-4, 10.0, 200), [1 / 336.0, 1 / 1.6, 1.0]]):
            hi, lo = time_scale_dd(c, float(tau), name)
            exact = abs(mp.mpf(float(tau))) * DD_VALUES[name]()
            worst_scale = max(worst_scale, float(abs(mp.mpf(hi) + mp.mpf(lo) - exact) / exact))
            d = float(rng.uniform(0.1, 100.0))
            # scale_by: fma(v, s.x, v * s.y)
            vs = fma(d, hi, float(mp.mpf(d) * mp.mpf(lo)))
            m = float(mp.mpf(vs) * mp.mpf(vs))
            tau2 = float(mp.mpf(float(tau)) * mp.mpf(float(tau)))
            # from_base2_sq(m, tau2, unit): fma(m, 2ln2_hi unit, m (2ln2_lo unit)) / tau2
            chi, clo = c["kTwoLn2Hi"] * unit, c["kTwoLn2Lo"] * unit
            back = float(mp.mpf(fma(m, chi, float(mp.mpf(m) * mp.mpf(clo)))) / mp.mpf(tau2))
            worst_back = max(worst_back, float(abs(mp.mpf(back) - mp.mpf(d) ** 2) / mp.mpf(d) ** 2))
        assert worst_scale < 1e-30, (name, worst_scale)
        assert worst_back < 8 * 2.0 ** -53, (name, worst_back)
