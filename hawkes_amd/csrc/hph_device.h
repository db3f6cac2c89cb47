// Device helpers shared by the pair kernels (gfx950, wave64).
#pragma once

#include "hph_exp_coeffs.h"
#include "hph_kernels.h"

namespace hph {
namespace kern {

// Excitation-term mode of a tile: absent, unmasked (every column strictly earlier), masked per pair.
enum { kSNone = 0, kSFull = 1, kSMixed = 2 };

// FP32-bulk mode: the background exponential is evaluated by one v_exp_f32 (2^y) on the
// FP64-formed exponent scaled by log2(e) and rounded to FP32; the result is widened back to FP64.
constexpr double kLog2e = 1.4426950408889634;
__device__ __forceinline__ float exp2_f32(float x) { return __builtin_amdgcn_exp2f(x); }

// exp(x) in FP64 on the VALU (gfx950 has no FP64 transcendental instruction).
// Cody-Waite reduction with a shifter constant: k = rint(x log2 e) appears in the low word of
// fma(x, log2 e, 1.5*2^52); r = x - k ln2 in two fma; p(r) is the degree-10 polynomial of
// gen_exp_poly.py; 2^k is applied by v_ldexp_f64. The clamp keeps k inside int range for any
// finite x; for x < -1000 the result is 0 as in IEEE exp. Max relative error 3.4e-16 for x <= 709.
__device__ __forceinline__ double exp_fp64(double x) {
    x = fmax(x, -1000.0);
    const double shifter = 6755399441055744.0; // 1.5 * 2^52
    const double kf = __builtin_fma(x, HPH_LOG2E, shifter);
    const double k = kf - shifter;
    const int ki = (int)(unsigned)__builtin_bit_cast(unsigned long long, kf);
    double r = __builtin_fma(-k, HPH_LN2_HI, x);
    r = __builtin_fma(-k, HPH_LN2_LO, r);
    double p = HPH_EXP_C10;
    p = __builtin_fma(p, r, HPH_EXP_C9);
    p = __builtin_fma(p, r, HPH_EXP_C8);
    p = __builtin_fma(p, r, HPH_EXP_C7);
    p = __builtin_fma(p, r, HPH_EXP_C6);
    p = __builtin_fma(p, r, HPH_EXP_C5);
    p = __builtin_fma(p, r, HPH_EXP_C4);
    p = __builtin_fma(p, r, HPH_EXP_C3);
    p = __builtin_fma(p, r, HPH_EXP_C2);
    p = __builtin_fma(p, r, HPH_EXP_C1);
    p = __builtin_fma(p, r, HPH_EXP_C0);
    return __builtin_amdgcn_ldexp(p, ki);
}

// 2^y in FP64 for |y| < 2^31, without a clamp (callers use it only where the host has bounded
// every exponent of the pass below 1e9 in size): k = rint(y) appears in the low word of
// y + 1.5*2^52, f = y - k is exact (|f| <= 1/2), q(f) is the degree-10 polynomial of
// gen_exp_poly.py and v_ldexp_f64 applies 2^k. One operation fewer than exp_fp64 (no two-part
// ln2 reduction); max relative error 3.3e-16; exactly 0 for y < -1075.
__device__ __forceinline__ double exp2_fp64(double y) {
    const double shifter = 6755399441055744.0; // 1.5 * 2^52
    const double kf = y + shifter;
    const double k = kf - shifter;
    const int ki = (int)(unsigned)__builtin_bit_cast(unsigned long long, kf);
    const double f = y - k;
    double p = HPH_EXP2_C10;
    p = __builtin_fma(p, f, HPH_EXP2_C9);
    p = __builtin_fma(p, f, HPH_EXP2_C8);
    p = __builtin_fma(p, f, HPH_EXP2_C7);
    p = __builtin_fma(p, f, HPH_EXP2_C6);
    p = __builtin_fma(p, f, HPH_EXP2_C5);
    p = __builtin_fma(p, f, HPH_EXP2_C4);
    p = __builtin_fma(p, f, HPH_EXP2_C3);
    p = __builtin_fma(p, f, HPH_EXP2_C2);
    p = __builtin_fma(p, f, HPH_EXP2_C1);
    p = __builtin_fma(p, f, HPH_EXP2_C0);
    return __builtin_amdgcn_ldexp(p, ki);
}

// 2^-y for 0 <= y < 2^31 (the background exponents of the base-2 form, bounded below 1e9 by the
// host as for exp2_fp64): y = i + f with i = trunc(y) (v_cvt_i32_f64 of -y gives -i) and
// f = fract(y) in [0, 1) (v_fract_f64, exact for y >= 0); s(f) is the degree-10 polynomial of
// gen_exp_poly.py for 2^-f on [0, 1] and v_ldexp_f64 applies 2^-i. Two operations for the
// reduction where exp2_fp64 takes three; max relative error 5.0e-16; exactly 0 for y > 1075.
__device__ __forceinline__ double exp2_neg_fp64(double y) {
    const int mi = (int)(-y);
    const double f = __builtin_amdgcn_fract(y);
    double p = HPH_EXP2N_C10;
    p = __builtin_fma(p, f, HPH_EXP2N_C9);
    p = __builtin_fma(p, f, HPH_EXP2N_C8);
    p = __builtin_fma(p, f, HPH_EXP2N_C7);
    p = __builtin_fma(p, f, HPH_EXP2N_C6);
    p = __builtin_fma(p, f, HPH_EXP2N_C5);
    p = __builtin_fma(p, f, HPH_EXP2N_C4);
    p = __builtin_fma(p, f, HPH_EXP2N_C3);
    p = __builtin_fma(p, f, HPH_EXP2N_C2);
    p = __builtin_fma(p, f, HPH_EXP2N_C1);
    p = __builtin_fma(p, f, HPH_EXP2N_C0);
    return __builtin_amdgcn_ldexp(p, mi);
}

// 2^-y for y >= 0 from a table of S = 256 entries: the FP64 background terms of the base-2 form,
// whose exponent arrives scaled by S (yS = S y, a sum of squares of coordinates and times scaled
// by time_scale_tab_dd; the host keeps yS below 2^31). -trunc(yS) = S a + b with a = its
// arithmetic shift by 8 and b in [0, S - 1] its low byte, f = fract(yS) in [0, 1) (exact for
// yS >= 0):
//   2^-y = 2^a tab[b] u(f),   tab[b] = 2^(b/S),   u(f) = 1 + f q(f) ~ 2^(-f/S)
// with u the degree-4 polynomial of gen_exp_poly.py, applied as tab[b] + tab[b] (f q(f)), and
// v_ldexp_f64 for 2^a. Ten VALU operations (convert, fract, the byte address - one shift whose
// operand is the low byte, an SDWA select -, shift, 3 fma, mul, fma, ldexp) and one LDS read where
// exp2_neg_fp64 takes 13. The table is 2 KB. Max relative error 2.0e-16 (table entries correctly
// rounded, so their errors do not share a sign as a polynomial's do); exactly 0 below the FP64
// range, as IEEE exp.
constexpr int kExp2TabBits = HPH_EXP2T_BITS;
constexpr int kExp2TabSize = 1 << kExp2TabBits;
static_assert(kExp2TabBits == 8, "the byte index (ni & 255, one SDWA operand select) and the scale constants");
// f q(f), q = HPH_EXP2T_Q by Horner
__device__ __forceinline__ double exp2t_fq(double f) {
    constexpr double Q[] = {HPH_EXP2T_Q};
    constexpr int n = (int)(sizeof(Q) / sizeof(Q[0]));
    double q = Q[n - 1];
#pragma unroll
    for (int k = n - 2; k >= 0; --k) q = __builtin_fma(q, f, Q[k]);
    return f * q;
}
// The exponential in two halves, so that a pass can issue the table read of one term and finish
// another meanwhile: exp2_tab_split reduces yS (ni = -trunc(yS), f = fract(yS)), exp2_tab_entry
// reads tab[b], exp2_tab_finish forms the result (KT: exp2_neg_ktab's table and bound, below).
struct Exp2TabArg {
    int ni;
    double f;
};
__device__ __forceinline__ Exp2TabArg exp2_tab_split(double yS) { return {(int)(-yS), __builtin_amdgcn_fract(yS)}; }
__device__ __forceinline__ double exp2_tab_entry(const double *tab, int ni) { return tab[ni & (kExp2TabSize - 1)]; }
template <bool KT> __device__ __forceinline__ double exp2_tab_finish(double t, Exp2TabArg a) {
    const double fq = exp2t_fq(a.f);
    if (KT) {
        uint2 w = __builtin_bit_cast(uint2, t);
        w.y += (unsigned)a.ni << (20 - kExp2TabBits);
        t = __builtin_bit_cast(double, w);
        return __builtin_fma(t, fq, t);
    }
    return __builtin_amdgcn_ldexp(__builtin_fma(t, fq, t), a.ni >> kExp2TabBits);
}
__device__ __forceinline__ double exp2_neg_tab(double yS, const double *tab) {
    const Exp2TabArg a = exp2_tab_split(yS);
    return exp2_tab_finish<false>(exp2_tab_entry(tab, a.ni), a);
}
// exp2_neg_tab where every argument keeps the result in the normal range with margin
// (0 <= yS <= kExp2KTabMaxY, checked by the caller for all pairs of a launch): the power 2^k,
// k = ni >> 8, goes into the table entry's exponent field by one integer add instead of a shift
// and an ldexp after the fma. ni = 256 k + b, so adding ni << 12 to the high word adds k << 20 (the
// exponent) and b << 12; the table (fill_exp2_table, ktab) holds 2^(b/256) with b << 12 taken off
// its high word, so the sum is exactly the bit pattern of 2^(b/256) 2^k. With k >= -1011 every
// t 2^k and the fma's result are normal doubles, where scaling by 2^k commutes with rounding: the
// result is bitwise exp2_neg_tab's.
constexpr double kExp2KTabMaxY = 258816.0; // 256 x 1011: k = ni >> 8 >= -1011
static_assert(kExp2KTabMaxY == kExp2TabSize * 1011.0, "k >= -1011");
__device__ __forceinline__ double exp2_neg_ktab(double yS, const double *tab) {
    const Exp2TabArg a = exp2_tab_split(yS);
    return exp2_tab_finish<true>(exp2_tab_entry(tab, a.ni), a);
}
// exp2_neg_tab's result from exp2_neg_ktab's table, for any yS >= 0 (also past exp2_neg_ktab's
// range): the entry's b << 12, taken off its high word in that table, is added back first (two
// integer operations), then the ldexp form. Kernels that hold the ktab table use it for the
// terms whose exponents the host does not bound (the excitation terms).
__device__ __forceinline__ double exp2_neg_tab_on_ktab(double yS, const double *tab) {
    const Exp2TabArg a = exp2_tab_split(yS);
    uint2 w = __builtin_bit_cast(uint2, exp2_tab_entry(tab, a.ni));
    w.y += (unsigned)(a.ni & (kExp2TabSize - 1)) << (20 - kExp2TabBits);
    return exp2_tab_finish<false>(__builtin_bit_cast(double, w), a);
}
// Fills exp2_neg_tab's table (LDS): entry b = 2^(b/256) correctly rounded (gen_exp_poly.py); with
// ktab, exp2_neg_ktab's: the same entry with b << 12 taken off its high word (a bit pattern, not
// used as a number). Threads tid < 256 of the workgroup write it (a workgroup of 256 threads
// writes every entry; smaller ones loop); the caller barriers before use.
__device__ __forceinline__ void fill_exp2_table(double *tab, int tid, int nthreads, bool ktab = false) {
    constexpr double M[kExp2TabSize] = {HPH_EXP2T_TABLE};
    for (int b = tid; b < kExp2TabSize; b += nthreads) {
        uint2 w = __builtin_bit_cast(uint2, M[b]);
        if (ktab) w.y -= (unsigned)b << (20 - kExp2TabBits);
        tab[b] = __builtin_bit_cast(double, w);
    }
}

// Scaled time of the base-2 pair form: t' = (t - t0) * st with st = |tauT| sqrt(log2(e) / 2), so
// that the background exponent tauT^2 dt^2 / 2 becomes dt'^2 / log2(e) and 2^-(dt'^2) =
// e^-(tauT^2 dt^2 / 2). Kernels and finalize compute st with this one expression.
constexpr double kSqrtHalfLog2e = 0.84932180028801904272; // sqrt(log2(e) / 2)
__device__ __forceinline__ double time_scale(double tauT) { return fabs(tauT) * kSqrtHalfLog2e; }

// Double-double constants (hi + lo, |lo| <= ulp(hi) / 2; from a 50-digit evaluation, pinned by
// tests/test_exp_poly.py): sqrt(log2(e) / 2), 2 ln 2 = 1 / (sqrt(log2(e) / 2))^2 and log2(e).
constexpr double kSqrtHalfLog2eHi = 0.8493218002880191, kSqrtHalfLog2eLo = -2.5959464104597403e-17;
constexpr double kTwoLn2Hi = 1.3862943611198906, kTwoLn2Lo = 4.638093627692599e-17;
constexpr double kLog2eHi = 1.4426950408889634, kLog2eLo = 2.0355273740931033e-17;
// sqrt(128 log2(e)) = sqrt(256) sqrt(log2(e) / 2): the scale of the table form (exponents times
// 256 = kExp2TabSize).
constexpr double kSqrt128Log2eHi = 13.589148804608305, kSqrt128Log2eLo = -4.1535142567355846e-16;

// The base-2 form's scales as double-doubles: a rounded scale st would make every scaled time
// st (t - t0) carry the same relative error, i.e. the pass would evaluate the background kernel at
// a slightly different tauT than the finalize's gradient formulas use; components of the gradient
// that cancel between events amplify that mismatch (measured: 1e-15 of the summation scale). With
// the scale carried to ~2^-104 and applied by one fma, only the random rounding of each scaled
// value remains. s = (hi, lo) with hi + lo = |tau| sqrt(log2(e) / 2).
__device__ __forceinline__ double2 time_scale_dd(double tau) {
    const double a = fabs(tau);
    const double hi = a * kSqrtHalfLog2eHi;
    const double err = __builtin_fma(a, kSqrtHalfLog2eHi, -hi); // exact error of the product
    return make_double2(hi, __builtin_fma(a, kSqrtHalfLog2eLo, err));
}
// The same with |tau| sqrt(128 log2(e)) (exponents scaled by 256, for exp2_neg_tab).
__device__ __forceinline__ double2 time_scale_tab_dd(double tau) {
    const double a = fabs(tau);
    const double hi = a * kSqrt128Log2eHi;
    const double err = __builtin_fma(a, kSqrt128Log2eHi, -hi);
    return make_double2(hi, __builtin_fma(a, kSqrt128Log2eLo, err));
}
// v (s.hi + s.lo), rounded once.
__device__ __forceinline__ double scale_by(double v, double2 s) { return __builtin_fma(v, s.x, v * s.y); }
// A moment summed in the base-2 form's squared units (r'^2 = r^2 tau^2 log2(e) / 2, or 256 times
// that in the table form: unit = 1/256) back to natural units: m unit 2 ln2 / tau2 with
// tau2 = tau * tau as the finalize's gradient formulas round it, so that tau2 times the result is
// m unit 2 ln 2 to one rounding. unit is a power of two (the constants' products are exact).
__device__ __forceinline__ double from_base2_sq(double m, double tau2, double unit = 1.0) {
    return __builtin_fma(m, kTwoLn2Hi * unit, m * (kTwoLn2Lo * unit)) / tau2;
}

// Background exponential exp(cBt dt^2 + cBx r^2): FP64 (exp_fp64) or FP32-bulk (v_exp_f32; the
// constants are then pre-scaled by log2 e).
template <bool F32>
__device__ __forceinline__ double background_exp(double cBx, double cBt, double r2, double dt2) {
    const double arg = __builtin_fma(cBt, dt2, cBx * r2);
    if (F32) return (double)exp2_f32((float)arg);
    return exp_fp64(arg);
}

// Next work item of a persistent pass (the same index in every thread of the workgroup). The
// queue counter is 0 when a launch starts and the launch returns it to 0: each workgroup fetches
// until it receives an index >= L.items, so a launch makes exactly L.items + L.grid fetches, and
// the workgroup that receives the last index, L.items + L.grid - 1 (every other fetch of the launch
// precedes it on the counter), stores 0. The next launch on the stream then finds it at 0
// without a reset kernel. L.grid must be the launched grid.
__device__ __forceinline__ int queue_next(const SymLaunch &L, int *s_item) {
    if (threadIdx.x == 0) {
        const int v = atomicAdd(L.counter, 1);
        if (v == L.items + L.grid - 1) __hip_atomic_store(L.counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_item = v;
    }
    __syncthreads();
    const int idx = __builtin_amdgcn_readfirstlane(*s_item); // uniform: item data in scalar registers
    __syncthreads();
    return idx;
}

// Stride (doubles) of one event in an LDS tile: D coordinates then t, padded to 16 bytes.
template <int D> struct TileStride {
    static constexpr int value = (D + 2) & ~1;
};

// Stride of one event in a location-gradient tile: D coordinates, t, w = 1/lambda, padded.
template <int D> struct LocStride {
    static constexpr int value = (D + 3) & ~1;
};

// Wavefronts per workgroup of the symmetric passes.
constexpr int kSymWaves = 4;

// Value of lane (byte_addr / 4) of the wavefront (ds_bpermute through the LDS crossbar).
__device__ __forceinline__ double lane_shift_down(double v, int byte_addr) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)u);
    const int hi = __builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// ---- double-double helpers (Knuth TwoSum) ----
__device__ __forceinline__ double2 dd_add(double2 a, double2 b) {
    const double s = a.x + b.x;
    const double bb = s - a.x;
    double e = (a.x - (s - bb)) + (b.x - bb);
    e += a.y + b.y;
    const double hi = s + e;
    return make_double2(hi, e - (hi - s));
}

__device__ __forceinline__ double ipow(double x, int e) {
    double r = 1.0;
    for (int q = 0; q < e; ++q) r *= x;
    return r;
}

// Per-row log-likelihood term and gradient terms from the row's exp sums (src/NewHawkes.hpp:652-655):
//   ll_i = log(lambda_i) + theta (e^{-omega (T - t_i)} - 1) - mu0 (Phi(tauT (T - t_i)) - Phi(-tauT t_i))
// s holds (sum eB, sum eB r^2, sum eB dt^2, sum eS, sum eS r^2, sum eS dt) with
// eB = exp(-(tauX^2 r^2 + tauT^2 dt^2)/2) = 2 pi pdf(tauX r) pdf(tauT dt) and
// eS = exp(-omega dt - sigma^2 r^2 / 2) = sqrt(2 pi) e^{-omega dt} pdf(sigma r). Without GRAD only
// s[0] and s[3] are used. v receives ll_i and d ll_i / d(sigma, tauX, tauT, omega, theta, mu0),
// lam the rate lambda_i; returns the self-excitation probability p_i.
template <bool GRAD>
__device__ __forceinline__ double row_terms(const double (&s)[6], const double *__restrict__ pk, int D, double ti,
                                            double t_last, double (&v)[7], double &lam_out) {
        const double sB = s[0], sS = s[3];
        const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3], theta = pk[4], mu0 = pk[5];
        const double inv_2pi = 0.15915494309189533577;
        const double inv_sqrt_2pi = 0.39894228040143267794;
        const double sigmaD = ipow(sigma, D), tauXD = ipow(tauX, D);
        const double bg = mu0 * tauXD * tauT * inv_2pi * sB;
        const double se = theta * omega * sigmaD * inv_sqrt_2pi * sS;
        const double lam = bg + se;
        lam_out = lam;
        const double dT = t_last - ti;
        const double eT = exp(-omega * dT);
        const double phiA = 0.5 * erfc(-tauT * dT * M_SQRT1_2);
        const double phiB = 0.5 * erfc(tauT * ti * M_SQRT1_2);
        v[0] = log(lam) + theta * (eT - 1.0) - mu0 * (phiA - phiB);
        if (GRAD) {
            const double pB = sB * inv_2pi, pBr = s[1] * inv_2pi, pBd = s[2] * inv_2pi;
            const double pS = sS * inv_sqrt_2pi, pSr = s[4] * inv_sqrt_2pi, pSd = s[5] * inv_sqrt_2pi;
            const double inv = 1.0 / lam;
            const double zA = tauT * dT, zB = tauT * ti;
            const double pdfA = exp(-0.5 * zA * zA) * inv_sqrt_2pi;
            const double pdfB = exp(-0.5 * zB * zB) * inv_sqrt_2pi;
            v[1] = theta * omega * ipow(sigma, D - 1) * (D * pS - sigma * sigma * pSr) * inv;
            v[2] = mu0 * tauT * ipow(tauX, D - 1) * (D * pB - tauX * tauX * pBr) * inv;
            v[3] = mu0 * tauXD * (pB - tauT * tauT * pBd) * inv - mu0 * (dT * pdfA + ti * pdfB);
            v[4] = theta * sigmaD * (pS - omega * pSd) * inv - theta * dT * eT;
            v[5] = omega * sigmaD * pS * inv + (eT - 1.0);
            v[6] = tauXD * tauT * pB * inv - (phiA - phiB);
        }
        return se / (bg + se);
}

// Fixed-order tree reduction of W double-doubles per thread over a 256-thread block, all W in
// one set of barriers: sh holds W x 256 entries; afterwards v holds the block sums (in every
// thread). The pairing is fixed, so results are bitwise reproducible.
template <int W>
__device__ void block_reduce_dd_w(double2 (&v)[W], double2 *sh) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < W; ++q) sh[q * kRowsPerBlock + tid] = v[q];
    __syncthreads();
    for (int s = kRowsPerBlock / 2; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int q = 0; q < W; ++q)
                sh[q * kRowsPerBlock + tid] = dd_add(sh[q * kRowsPerBlock + tid], sh[q * kRowsPerBlock + tid + s]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < W; ++q) v[q] = sh[q * kRowsPerBlock];
    __syncthreads();
}

// Writes the block's double-double sums of v (GRAD: all 7, else v[0]) to block_out[k][blk][q][2];
// sh holds 7 x 256 entries.
template <bool GRAD>
__device__ __forceinline__ void store_block_sums(const double (&v)[7], double2 *sh, double *__restrict__ block_out,
                                                 int k, int nblocks, int blk) {
    constexpr int W = GRAD ? 7 : 1;
    double2 r[W];
#pragma unroll
    for (int q = 0; q < W; ++q) r[q] = make_double2(v[q], 0.0);
    block_reduce_dd_w<W>(r, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            double *o = block_out + (((size_t)k * nblocks + blk) * kResultWidth + q) * 2;
            o[0] = r[q].x;
            o[1] = r[q].y;
        }
    }
}

} // namespace kern
} // namespace hph
