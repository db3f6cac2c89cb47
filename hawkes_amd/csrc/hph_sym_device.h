// Device code of the symmetric pair pass (included by hph_sym.hip and hph_split.hip).
//
// Symmetric pair pass: each background exp is formed once per unordered pair.
//
// The background kernel B_ij = pdf(tauX r_ij) pdf(tauT dt_ij) is symmetric in (i, j), and so are
// its gradient moments B r^2 and B dt^2. With events time-sorted, a row block P (R rows: 4
// wavefronts x 64 lanes x A rows per lane) is paired only with earlier columns (j < P*R):
//   * row side:    B, B r^2, B dt^2 and the excitation sums S, S r^2, S dt for the block's rows
//                  (all excitation of row i comes from earlier events);
//   * column side: B, B r^2, B dt^2 for the earlier events j, written per (block, column) and
//                  summed over blocks in the finalize kernel in a fixed order.
// The block's own columns (the diagonal block) are paired in both orders with the per-pair
// dt > 0 mask, as in src/NewHawkes.hpp:520-547.
//
// Column sums inside a wavefront: over 64 steps lane l pairs its rows with column (l + s) mod 64
// of a 64-column subtile, sums its rows' contributions to that column and adds them to slot
// l + s of 128 LDS slots (ds_add_f64: no result to wait for; in one step the 64 lanes add to 64
// different slots, so every slot receives its additions in step order). Column u's sum is then
// slot u (steps s <= u) plus slot u + 64 (the steps after the wrap). Each wave has its own slots,
// so the sums do not depend on the timing of the waves. The subtile is stored twice in LDS, as
// structure of arrays, so the per-lane read address (l + s) needs no wrap either and
// consecutive lanes read consecutive doubles (no bank conflicts).
#pragma once

#include "hph_device.h"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace hph {
namespace kern {


// Per-vector constants of the pair loops, in one of two forms.
//  natural (FAST false): coordinates and times as given; exponents in e-units with the clamped
//      exp_fp64: background cB0 r^2 + cB1 dt^2 (cB0 = -tauX^2/2, cB1 = -tauT^2/2; times log2 e
//      for the FP32-bulk v_exp_f32), excitation cS0 r^2 + cS1 dt (cS0 = -sigma^2/2, cS1 = -omega).
//  base 2 (FAST true): coordinates arrive as x' = sx (x - x0) and times as t' = st (t - t0)
//      (time_scale_dd of tauX and tauT, applied by one fma each; x0, t0 = the row block's first
//      event), so that the background exponent in 2-units is the plain sum of squares y = r'^2 + dt'^2
//      (clamp-free exp2_neg_fp64); excitation -(cS0 r'^2 + cS1 dt) / 256 in 2-units, with
//      cS0 = 256 sigma^2 / tauX^2, cS1 = 256 omega log2(e) (double-doubles) and dt the difference
//      of the natural times (staged beside the scaled ones: a scaled difference carries the
//      rounding of st (t - t0), which an excitation term would feel as a relative error of order
//      eps omega |t - t0|), evaluated by exp2_neg_tab (excitation_term). The background moments
//      are summed as e y and e dt'^2, the excitation ones as e r'^2 and e dt; the finalize kernels
//      turn them into e r^2 = (e y - e dt'^2) / sx^2, e dt^2, e r^2 and e dt.
//      In FP64 (F32 false) the scales are sqrt(256) times larger (time_scale_tab_dd), so that y
//      arrives as 256 times the exponent for the table exponential exp2_neg_tab, and the finalize
//      divides the squared moments by 256 (SymLaunch::b2tab); cS0 is then sigma^2 / tauX^2.
struct PairConst {
    double cB0, cB1, cS0, cS1;
    double cS0lo, cS1lo; // base 2: the low parts of cS0 and cS1 (double-doubles), else 0
    const double *tab;   // base 2: exp2_neg_tab's table (LDS), else unused
};

// Pair geometry from the row's coordinates xi, ti and the column's q (D coordinates, then the
// time): dt, dt^2 and, in the base-2 form, the background exponent y = r'^2 + dt'^2. r^2 is formed
// where the natural form or the excitation term needs it (R2); else y takes the squares directly.
template <int D, bool FAST, bool R2, int NQ>
__device__ __forceinline__ void pair_geometry(const double (&xi)[D], double ti, const double (&q)[NQ], double &r2,
                                              double &dt, double &dt2, double &y) {
    static_assert(NQ >= D + 1, "a column holds D coordinates and the time");
    dt = ti - q[D];
    dt2 = dt * dt;
    if (FAST && !R2) {
        y = dt2;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double dd = xi[d] - q[d];
            y = __builtin_fma(dd, dd, y);
        }
        r2 = 0.0;
        return;
    }
    {
        const double d0 = xi[0] - q[0];
        r2 = d0 * d0;
    }
#pragma unroll
    for (int d = 1; d < D; ++d) {
        const double dd = xi[d] - q[d];
        r2 = __builtin_fma(dd, dd, r2);
    }
    y = FAST ? r2 + dt2 : 0.0;
}

// Background term of one pair (geometry from pair_geometry). KT: the launch's exponents all lie
// in exp2_neg_ktab's range (SymLaunch::ktab; the table is exp2_neg_ktab's).
template <bool F32, bool FAST, bool KT = false>
__device__ __forceinline__ double background_term(const PairConst &c, double r2, double dt2, double y,
                                                  const double *__restrict__ tab) {
    if (FAST) {
        if (F32) return (double)exp2_f32((float)(-y));
        // y = 256 x the exponent >= 0: a sum of squares
        return KT ? exp2_neg_ktab(y, tab) : exp2_neg_tab(y, tab);
    }
    return background_exp<F32>(c.cB0, c.cB1, r2, dt2);
}

// Excitation term of one pair. Base 2 (double-double constants: the exponent's scale matches the
// finalize's sigma and omega; TAB in both precisions): forms yS = cS0 r'^2 + cS1 dt = -256 x the
// base-2 exponent and evaluates exp2_neg_tab(yS), 2^(-yS/256) (KT: the table is exp2_neg_ktab's,
// whose range bounds only the background exponents: exp2_neg_tab_on_ktab). For a causal pair yS >= 0; pairs with dt <= 0,
// whose yS may be negative (the result then finite or +inf, not 2^(-yS/256)), are masked by a
// select by every caller; yS past 2^31 (terms far below the FP64 range) converts to INT_MIN and
// gives +0.
template <bool FAST, bool TAB, bool KT>
__device__ __forceinline__ double excitation_term(const PairConst &c, double r2, double dt,
                                                  const double *__restrict__ tab) {
    if (FAST) {
        const double u = __builtin_fma(c.cS1, dt, c.cS1lo * dt);
        const double y = __builtin_fma(c.cS0, r2, __builtin_fma(c.cS0lo, r2, u));
        if (!TAB) return exp2_fp64(y);
        return KT ? exp2_neg_tab_on_ktab(y, tab) : exp2_neg_tab(y, tab);
    }
    return exp_fp64(__builtin_fma(c.cS0, r2, c.cS1 * dt));
}

// Adds v to an LDS double (ds_add_f64; the result is not read).
__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Doubles of one staged 64-column subtile: the D coordinates and the time as rows of 128 (each
// 64-column run stored twice, so that the per-lane read index lane + s needs no wrap) and, in the
// base-2 form, the natural time as one row of 64, read only by subtiles that form excitation
// terms (a small share of the pass), with a wrapped index.
template <int D, bool FAST> struct SubStride {
    static constexpr int value = (D + 1) * 128 + (FAST ? 64 : 0);
};

// Background-only subtile in the FP64 table form (the bulk of a likelihood pass), pipelined by one
// step: step s + 1's geometry, its reduction and its table reads are issued before step s's
// exponential is finished, so each term's LDS read waits out a step of arithmetic instead of
// standing in its term's dependency chain. The pairs, the operations on each and the order of
// every addition are those of sym_subtile: the results are bitwise the same.
template <int D, bool GRAD, int A, bool TAIL, bool KT>
__device__ __forceinline__ void sym_subtile_bg_tab(const double *__restrict__ sub, int lane, const double (&xi)[A][D],
                                                   const double (&ti)[A], const double (&vm)[A], double (&acc)[A][6],
                                                   double *cs, const double *__restrict__ tab) {
    constexpr int NC = GRAD ? 3 : 1;
    constexpr int NQ = D + 1;
    // one term in flight: y, dt'^2, the reduced exponent and the table entry
    struct Term {
        double y, dt2, t;
        Exp2TabArg a;
    };
    auto issue = [&](const double (&q)[NQ], Term (&p)[A]) {
#pragma unroll
        for (int a = 0; a < A; ++a) {
            double r2, dt, dt2, y;
            pair_geometry<D, true, false>(xi[a], ti[a], q, r2, dt, dt2, y);
            p[a].y = y;
            p[a].dt2 = dt2;
            p[a].a = exp2_tab_split(y);
            p[a].t = exp2_tab_entry(tab, p[a].a.ni);
        }
    };
    Term cur[A];
    double qn[NQ];
    {
        double q[NQ];
#pragma unroll
        for (int d = 0; d < NQ; ++d) {
            q[d] = sub[d * 128 + lane];
            qn[d] = sub[d * 128 + lane + 1];
        }
        issue(q, cur);
    }
    // step s: columns lane + s (cur, issued) and lane + s + 1 (qn, read). Every step issues the
    // next one, the last step too: its read-ahead entry, lane + 65 <= 128, is the first of the
    // subtile's next row (inside the staged tile) and its term is discarded, which costs one step's
    // geometry per subtile and keeps the loop free of a branch.
    auto step = [&](int s) {
        Term nxt[A];
        {
            double q[NQ];
#pragma unroll
            for (int d = 0; d < NQ; ++d) {
                q[d] = qn[d];
                qn[d] = sub[d * 128 + lane + s + 2];
            }
            issue(q, nxt);
        }
        double v[NC];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const double eb = exp2_tab_finish<KT>(cur[a].t, cur[a].a);
            acc[a][0] += eb;
            if (GRAD) {
                acc[a][1] = __builtin_fma(eb, cur[a].y, acc[a][1]);
                acc[a][2] = __builtin_fma(eb, cur[a].dt2, acc[a][2]);
            }
            const double ec = TAIL ? eb * vm[a] : eb;
            if (a == 0) {
                v[0] = ec;
                if (GRAD) {
                    v[1] = ec * cur[a].y;
                    v[2] = ec * cur[a].dt2;
                }
            } else {
                v[0] += ec;
                if (GRAD) {
                    v[1] = __builtin_fma(ec, cur[a].y, v[1]);
                    v[2] = __builtin_fma(ec, cur[a].dt2, v[2]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) lds_add(cs + c * 128 + lane + s, v[c]);
#pragma unroll
        for (int a = 0; a < A; ++a) cur[a] = nxt[a];
    };
#pragma unroll 8
    for (int s = 0; s < 64; ++s) step(s);
}

// One 64-column subtile against this lane's A rows; the column contributions are added to the
// wave's LDS slots cs (moment c, slot u in [0, 128) at cs[c * 128 + u]).
// DOB: form the background term; SMODE: excitation absent / unmasked / masked; TAIL: the block
// has rows past the last event, whose column contributions are masked by vm; F32: FP32-bulk
// background exponential; FAST: the base-2 form of PairConst.
template <int D, bool GRAD, int A, bool DOB, int SMODE, bool TAIL, bool F32, bool FAST, bool KT>
__device__ __forceinline__ void sym_subtile(const double *__restrict__ sub, int lane, const double (&xi)[A][D],
                                            const double (&ti)[A], const double (&tn)[A], const double (&vm)[A],
                                            const PairConst &c, double (&acc)[A][6], double *cs,
                                            const double *__restrict__ tab) {
    // tab: exp2_neg_tab's table (c.tab); restrict: the column-slot additions to the same LDS
    // array never touch it
    constexpr int NC = GRAD ? 3 : 1;
    // components read per column: D coordinates and the time, plus the natural time where the
    // base-2 form forms excitation terms
    constexpr int NQ = (FAST && SMODE != kSNone) ? D + 2 : D + 1;
    // structure-of-arrays subtile (SubStride): component d <= D of entry u at sub[d * 128 + u],
    // the natural time at sub[(D + 1) * 128 + (u & 63)]; the next step's entry is read before this
    // step's slot additions are issued (entry lane + 64 of the last step's read is in the stored
    // copy, or wraps)
    double qn[NQ];
#pragma unroll
    for (int d = 0; d < NQ; ++d) qn[d] = sub[d * 128 + lane];
    auto step = [&](int s) {
        double q[NQ];
#pragma unroll
        for (int d = 0; d < NQ; ++d) {
            q[d] = qn[d];
            qn[d] = d <= D ? sub[d * 128 + lane + s + 1] : sub[d * 128 + ((lane + s + 1) & 63)];
        }
        double v[NC]; // this step's column contributions of the A rows
#pragma unroll
        for (int a = 0; a < A; ++a) {
            double r2, dt, dt2, y;
            pair_geometry<D, FAST, !FAST || SMODE != kSNone>(xi[a], ti[a], q, r2, dt, dt2, y);
            const double m1 = FAST ? y : r2; // first background moment: e y (base 2) or e r^2
            // the excitation term's time difference: natural in both forms
            double dte = dt;
            if constexpr (FAST && SMODE != kSNone) dte = tn[a] - q[D + 1];
            if (DOB) {
                const double eb = background_term<F32, FAST, KT>(c, r2, dt2, y, tab);
                acc[a][0] += eb;
                if (GRAD) {
                    acc[a][1] = __builtin_fma(eb, m1, acc[a][1]);
                    acc[a][2] = __builtin_fma(eb, dt2, acc[a][2]);
                }
                const double ec = TAIL ? eb * vm[a] : eb;
                if (a == 0) {
                    v[0] = ec;
                    if (GRAD) {
                        v[1] = ec * m1;
                        v[2] = ec * dt2;
                    }
                } else {
                    v[0] += ec;
                    if (GRAD) {
                        v[1] = __builtin_fma(ec, m1, v[1]);
                        v[2] = __builtin_fma(ec, dt2, v[2]);
                    }
                }
            }
            if (SMODE != kSNone) {
                double es = excitation_term<FAST, FAST, KT>(c, r2, dte, tab);
                if (SMODE == kSMixed) es = dte > 0.0 ? es : 0.0;
                acc[a][3] += es;
                if (GRAD) {
                    acc[a][4] = __builtin_fma(es, r2, acc[a][4]);
                    acc[a][5] = __builtin_fma(es, dte, acc[a][5]);
                }
            }
        }
        if (DOB) {
#pragma unroll
            for (int c = 0; c < NC; ++c) lds_add(cs + c * 128 + lane + s, v[c]);
        }
    };
    // the background-only loop (the bulk of a likelihood pass) is unrolled further: its
    // per-iteration address and constant moves are then shared by more pairs
    if constexpr (DOB && SMODE == kSNone && FAST && !F32) {
        sym_subtile_bg_tab<D, GRAD, A, TAIL, KT>(sub, lane, xi, ti, vm, acc, cs, tab);
    } else if constexpr (DOB && SMODE == kSNone) {
#pragma unroll 8
        for (int s = 0; s < 64; ++s) step(s);
    } else {
#pragma unroll 2
        for (int s = 0; s < 64; ++s) step(s);
    }
}

// Stride (doubles) of one event in a diagonal-block tile: D coordinates, the time and, in the
// base-2 form, the natural time, padded to 16 bytes.
template <int D, bool FAST> struct DiagStride {
    static constexpr int value = FAST ? LocStride<D>::value : TileStride<D>::value;
};

// One diagonal-block tile (plain layout, broadcast reads) against one row: ordered pairs with
// the dt > 0 mask. PM: the terms of the pass mode (kPassBoth, or kPassCulledB: background only).
// tn: the row's natural time (the base-2 form's excitation terms).
template <int D, bool GRAD, bool F32, bool FAST, bool KT, int PM>
__device__ __forceinline__ void diag_pass(const double *__restrict__ tile, int cnt, const double (&xi)[D], double ti,
                                          double tn, const PairConst &c, double (&acc)[6]) {
    constexpr int ST = DiagStride<D, FAST>::value;
    constexpr int NQ = (FAST && PM != kPassCulledB) ? D + 2 : D + 1;
#pragma unroll 2
    for (int j = 0; j < cnt; ++j) {
        double q[NQ];
#pragma unroll
        for (int d = 0; d < NQ; ++d) q[d] = tile[j * ST + d];
        double r2, dt, dt2, y;
        pair_geometry<D, FAST, !FAST || PM != kPassCulledB>(xi, ti, q, r2, dt, dt2, y);
        double dte = dt;
        if constexpr (FAST && PM != kPassCulledB) dte = tn - q[D + 1];
        {
            const double eb = background_term<F32, FAST, KT>(c, r2, dt2, y, c.tab);
            acc[0] += eb;
            if (GRAD) {
                acc[1] = __builtin_fma(eb, FAST ? y : r2, acc[1]);
                acc[2] = __builtin_fma(eb, dt2, acc[2]);
            }
        }
        if (PM != kPassCulledB) {
            double es = excitation_term<FAST, FAST, KT>(c, r2, dte, c.tab);
            es = dte > 0.0 ? es : 0.0;
            acc[3] += es;
            if (GRAD) {
                acc[4] = __builtin_fma(es, r2, acc[4]);
                acc[5] = __builtin_fma(es, dte, acc[5]);
            }
        }
    }
}

// mode = 4 * (background formed) + excitation mode; the culled background pass builds only its own.
template <int D, bool GRAD, int A, bool TAIL, bool F32, bool FAST, bool KT, int PM>
__device__ __forceinline__ void sym_dispatch(int mode, const double *__restrict__ sub, int lane,
                                             const double (&xi)[A][D], const double (&ti)[A], const double (&tn)[A],
                                             const double (&vm)[A], const PairConst &c, double (&acc)[A][6],
                                             double *cs) {
    if constexpr (PM == kPassCulledB) {
        if (mode == 4 + kSNone) sym_subtile<D, GRAD, A, true, kSNone, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        return;
    }
    switch (mode) {
    case 4 + kSNone:
        sym_subtile<D, GRAD, A, true, kSNone, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        break;
    case 4 + kSFull:
        sym_subtile<D, GRAD, A, true, kSFull, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        break;
    case 4 + kSMixed:
        sym_subtile<D, GRAD, A, true, kSMixed, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        break;
    case kSFull:
        sym_subtile<D, GRAD, A, false, kSFull, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        break;
    case kSMixed:
        sym_subtile<D, GRAD, A, false, kSMixed, TAIL, F32, FAST, KT>(sub, lane, xi, ti, tn, vm, c, acc, cs, c.tab);
        break;
    default: break; // every term of the subtile is exactly zero
    }
}

// The culled pass's subtile against one of a lane's two rows (row 1 where s, else row 0; s is
// wave-uniform): the row's registers are moved into a one-row call and its sums moved back, so the
// other row's sums are left as they were.
template <int D, bool GRAD, bool F32, bool FAST, bool KT, int PM>
__device__ __forceinline__ void sym_one_row(bool s, bool tail, int mode, const double *__restrict__ sub, int lane,
                                            const double (&xi)[2][D], const double (&ti)[2], const double (&tn)[2],
                                            const double (&vm)[2], const PairConst &c, double (&acc)[2][6],
                                            double *cs) {
    constexpr int NB = GRAD ? 3 : 1; // the background moments: the only sums this pass forms
    double x1[1][D], t1[1], n1[1], v1[1], a1[1][6];
#pragma unroll
    for (int d = 0; d < D; ++d) x1[0][d] = s ? xi[1][d] : xi[0][d];
    t1[0] = s ? ti[1] : ti[0];
    n1[0] = s ? tn[1] : tn[0];
    v1[0] = s ? vm[1] : vm[0];
#pragma unroll
    for (int q = 0; q < 6; ++q) a1[0][q] = q < NB ? (s ? acc[1][q] : acc[0][q]) : 0.0;
    if (tail)
        sym_dispatch<D, GRAD, 1, true, F32, FAST, KT, PM>(mode, sub, lane, x1, t1, n1, v1, c, a1, cs);
    else
        sym_dispatch<D, GRAD, 1, false, F32, FAST, KT, PM>(mode, sub, lane, x1, t1, n1, v1, c, a1, cs);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        acc[0][q] = s ? acc[0][q] : a1[0][q];
        acc[1][q] = s ? a1[0][q] : acc[1][q];
    }
}

// Computed column tiles per partial row sum of a chunk item (sym_item).
constexpr int kFlushTiles = 4;

// Column side of the symmetric pass, read back per event: the later blocks Q whose pass wrote
// column tile t (flag set) are listed in ascending order, kQChunk at a time, so that the loads
// of the listed blocks are independent (no branch on a flag between them).
constexpr int kQChunk = 1024;

// Lists the blocks Q in [qa, qb) (qb - qa <= kQChunk) whose flag is set into list (ascending)
// and returns their count. ftile: the flags of column tile t, one byte per block Q (the
// [tile][block] layout of SymLaunch::colflag). Called by every thread of the block; wave 0 builds
// the list. The caller must barrier before the list is rebuilt. An entry is Q, with bit
// kListFull set where the flag is 2 (row slots whose excitation moments were written, sym_item).
constexpr int kListFull = 1 << 30;
__device__ __forceinline__ int flagged_blocks(const unsigned char *__restrict__ ftile, int qa, int qb, int *list,
                                              int *count) {
    if (threadIdx.x < 64) {
        int base = 0;
        const unsigned long long below = (1ULL << threadIdx.x) - 1ULL;
        for (int q0 = qa; q0 < qb; q0 += 64) {
            const int Q = q0 + (int)threadIdx.x;
            const int fl = Q < qb ? ftile[Q] : 0;
            const bool f = fl != 0;
            const unsigned long long m = __ballot(f);
            if (f) list[base + __popcll(m & below)] = fl >= 2 ? (Q | kListFull) : Q;
            base += __popcll(m);
        }
        if (threadIdx.x == 0) *count = base;
    }
    __syncthreads();
    return *count;
}

// Adds the column partials of event i of the listed blocks (ascending) to cs.
template <int NC>
__device__ __forceinline__ void add_listed_columns(const double *__restrict__ cbase, int R, int i, const int *list,
                                                   int cnt, double (&cs)[NC]) {
#pragma unroll 4
    for (int j = 0; j < cnt; ++j) {
        const int Q = list[j] & (kListFull - 1); // column flags are 0 or 1: no kListFull bit
        const double *cp = cbase + (size_t)NC * R * ((size_t)Q * (Q - 1) / 2);
        const size_t Q0 = (size_t)Q * R;
#pragma unroll
        for (int c = 0; c < NC; ++c) cs[c] += cp[(size_t)c * Q0 + i];
    }
}

// Column sums of the later blocks Q >= P + 1 + kQChunk (the first kQChunk later blocks are
// listed and added by the caller), ascending, skipping unflagged tiles. Every thread of the block
// calls it; rows past n (live false) only take part in the barriers.
template <int NC>
__device__ __forceinline__ void add_remaining_columns(const double *__restrict__ cbase,
                                                      const unsigned char *__restrict__ ftile, int R, int P, int nb,
                                                      int i, bool live, double (&cs)[NC], int *list, int *count) {
    for (int qa = P + 1 + kQChunk; qa < nb; qa += kQChunk) {
        __syncthreads(); // every thread has read the previous list
        const int cnt = flagged_blocks(ftile, qa, min(nb, qa + kQChunk), list, count);
        if (live) add_listed_columns<NC>(cbase, R, i, list, cnt, cs);
    }
}

// Row side of the symmetric pass, read back for row r of block P: adds the first NQ moments of
// the row partials of the chunk slots c < nch whose flag rf[c] is set (listed kQChunk at a time,
// ascending) and of the diagonal slot max_chunks into a. NR moments per slot (the first NR / 2
// of the background, then as many of the excitation), slots rpad rows apart. A chunk slot with
// flag 1 holds its background moments only: its excitation moments were exactly zero and were
// not written (adding them would change no bit). Every thread of the block calls it (rows past
// n: live false, barriers only); list and count are LDS scratch of the caller's own.
template <int NR, int NQ>
__device__ __forceinline__ void add_row_slots(const double *__restrict__ rbase, const unsigned char *__restrict__ rf,
                                              int nch, int max_chunks, size_t rpad, int r, bool live,
                                              double (&a)[NQ], int *list, int *count) {
    for (int c0 = 0; c0 < nch; c0 += kQChunk) {
        if (c0 > 0) __syncthreads(); // every thread has read the previous list
        const int cnt = flagged_blocks(rf, c0, min(nch, c0 + kQChunk), list, count);
        if (live) {
            // unrolled so that several slots' loads are in flight
#pragma unroll 4
            for (int u = 0; u < cnt; ++u) {
                const int e = list[u];
                const double *rp = rbase + (size_t)(e & (kListFull - 1)) * NR * rpad;
                const bool full = (e & kListFull) != 0;
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (q < NR / 2 || full) a[q] += rp[(size_t)q * rpad + r];
            }
        }
    }
    if (live) {
        const double *rp = rbase + (size_t)max_chunks * NR * rpad;
#pragma unroll
        for (int q = 0; q < NQ; ++q) a[q] += rp[(size_t)q * rpad + r];
    }
}

// Bounding box of `count` consecutive groups of the culled order, starting at group g:
// lo/hi[0..D-1] coordinates, lo/hi[D] time (gstat layout of SymData).
template <int D>
__device__ __forceinline__ void group_box(const double *__restrict__ gstat, int g, int count, double (&lo)[D + 1],
                                          double (&hi)[D + 1]) {
    constexpr int GS = 2 * D + 2;
#pragma unroll
    for (int d = 0; d <= D; ++d) {
        lo[d] = gstat[(size_t)g * GS + d];
        hi[d] = gstat[(size_t)g * GS + D + 1 + d];
    }
    for (int q = 1; q < count; ++q) {
#pragma unroll
        for (int d = 0; d <= D; ++d) {
            lo[d] = fmin(lo[d], gstat[(size_t)(g + q) * GS + d]);
            hi[d] = fmax(hi[d], gstat[(size_t)(g + q) * GS + D + 1 + d]);
        }
    }
}

// Smallest background exponent tauX^2 r^2 / 2 + tauT^2 dt^2 / 2 (natural units) over the pairs
// of two bounding boxes: the squared gaps between the boxes along every axis.
template <int D>
__device__ __forceinline__ double box_background_exponent(const double (&alo)[D + 1], const double (&ahi)[D + 1],
                                                          const double (&blo)[D + 1], const double (&bhi)[D + 1],
                                                          double hx, double ht) {
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double g = fmax(0.0, fmax(blo[d] - ahi[d], alo[d] - bhi[d]));
        r2 = __builtin_fma(g, g, r2);
    }
    const double gt = fmax(0.0, fmax(blo[D] - ahi[D], alo[D] - bhi[D]));
    return hx * r2 + ht * gt * gt;
}

// One work item of the symmetric pass: row block P of dataset it.x, vector it.w, against a chunk
// of earlier columns (it.z >= 0) or its own diagonal block (it.z < 0). FAST: the base-2 form.
// PM: pass mode (kPassBoth; kPassCulledB: background terms of the culled order, bounding-box
// skips, events of a chunk need not be earlier than the block's). FLUSH: chunks longer than
// kFlushTiles tiles sum their rows per group of kFlushTiles tiles (below); a separate variant,
// because the second set of row sums costs registers that short chunks do not need.
template <int D, bool GRAD, int A, bool F32, bool FAST, bool KT, int PM, bool FLUSH>
__device__ __forceinline__ void sym_item(const SymLaunch &L, const int4 it, const double *__restrict__ params,
                                         double *__restrict__ rowpart, double *__restrict__ colpart,
                                         double *__restrict__ tile, double *__restrict__ colbuf, double *cs,
                                         const double *tab, int tid, int lane, int w) {
    constexpr int ST = DiagStride<D, FAST>::value;
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    constexpr int R = kSymWaves * 64 * A;
    // doubles per staged chunk subtile: D coordinates, the time and, in the base-2 form, the
    // natural time
    constexpr int SUB = SubStride<D, FAST>::value;
    const SymData &ds = L.data[it.x];
    const int P = it.y, chunk = it.z, k = it.w;
    const double *pk = params + 6 * (size_t)(ds.param0 + k);
    const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3];
    const double *__restrict__ ex = ds.x;
    const double *__restrict__ et = ds.t;
    const int n = ds.n, npad = ds.npad;
    const int P0 = P * R;

    // base-2 form: coordinates and times relative to the block's first event, scaled by sx and st
    // (in FP64 sqrt(256) times larger: exponents times 256 for the table exponential)
    const double2 st = FAST ? (F32 ? time_scale_dd(tauT) : time_scale_tab_dd(tauT)) : make_double2(1.0, 0.0);
    const double2 sx = FAST ? (F32 ? time_scale_dd(tauX) : time_scale_tab_dd(tauX)) : make_double2(1.0, 0.0);
    const double t0 = FAST ? et[min(P0, n - 1)] : 0.0;
    double x0[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x0[d] = FAST ? ex[(size_t)d * npad + min(P0, n - 1)] : 0.0;
    PairConst c;
    c.cS0lo = 0.0;
    c.cS1lo = 0.0;
    c.tab = tab;
    if (FAST) {
        c.cB0 = 1.0;
        c.cB1 = 0.0;
        // -sigma^2 / tauX^2 with sigma^2 and tauX^2 rounded as the finalize rounds them, the
        // quotient carried as a double-double; -omega log2(e) per natural time difference, the
        // same. Both are multiplied by -256, exactly, so that the excitation exponent arrives as
        // yS = -256 x (its base-2 value) >= 0 for exp2_neg_tab (excitation_term); against the FP64
        // table form's r'^2 (256 times the base-2 units) the first becomes sigma^2 / tauX^2 itself
        const double s2 = sigma * sigma, x2 = tauX * tauX;
        const double q = s2 / x2;
        const double qlo = __builtin_fma(-q, x2, s2) / x2;
        const double w = -omega * kLog2eHi;
        const double wlo = __builtin_fma(-omega, kLog2eHi, -w) + (-omega * kLog2eLo);
        const double r2unit = F32 ? (double)kExp2TabSize : 1.0;
        c.cS0 = q * r2unit;
        c.cS0lo = qlo * r2unit;
        c.cS1 = -w * kExp2TabSize;
        c.cS1lo = -wlo * kExp2TabSize;
    } else {
        c.cB0 = -0.5 * tauX * tauX * (F32 ? kLog2e : 1.0);
        c.cB1 = -0.5 * tauT * tauT * (F32 ? kLog2e : 1.0);
        c.cS0 = -0.5 * sigma * sigma;
        c.cS1 = -omega;
    }
    // skip tests, on the natural times in both forms: a tile is skipped when the smallest time
    // gap alone drives every background exponent tauT^2 gap^2 / 2 past the negligible point or
    // every excitation exponent omega gap past the cut
    const double zb = ds.zero_arg_b;
    const double zs = L.cut_s ? L.xs[ds.param0 + k] : L.zero_arg_s; // (launch_sym_order's)
    const double gb = 0.5 * tauT * tauT; // background exponent per gap^2
    const double gs = omega;             // excitation exponent per gap
    // the lane's row a lies in 64-row group row_group(a) of the block: the wave's A groups are
    // consecutive, or in the culled order dealt round the waves (group a kSymWaves + w), since
    // groups near a column tile cluster along the curve and the tile's work is then spread over
    // the waves (the workgroup waits for its slowest wave at every tile)
    constexpr bool DEAL = PM == kPassCulledB;
    auto row_group = [w](int a) { return DEAL ? a * kSymWaves + w : w * A + a; };
    const int wrow0 = __builtin_amdgcn_readfirstlane(P0 + 64 * row_group(0)); // the wave's first row
    // tn: the rows' natural times (skip decisions, excitation terms); ti: their times in the
    // form's units (base 2: relative to the block's first event)
    double xi[A][D], ti[A], tn[A], vm[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int r = P0 + 64 * row_group(a) + lane;
        const int rc = min(r, n - 1);
#pragma unroll
        for (int d = 0; d < D; ++d) xi[a][d] = FAST ? scale_by(ex[(size_t)d * npad + rc] - x0[d], sx) : ex[(size_t)d * npad + rc];
        tn[a] = et[rc];
        ti[a] = FAST ? scale_by(tn[a] - t0, st) : tn[a];
        vm[a] = r < n ? 1.0 : 0.0;
    }
    const bool wave_live = wrow0 < n;
    const double t_lo = et[min(wrow0, n - 1)]; // the wave's earliest row (natural)

    double acc[A][6];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[a][q] = 0.0;

    int slot;
    bool full = true; // every moment of the row partials is written (diagonal items always)
    if (chunk < 0) {
        slot = ds.max_chunks;
        const int c1 = min(P0 + R, n);
        for (int j0 = P0; j0 < c1; j0 += kTile) {
            const int cnt = min(kTile, c1 - j0);
            __syncthreads();
            if (tid < cnt) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double v = ex[(size_t)d * npad + j0 + tid];
                    tile[tid * ST + d] = FAST ? scale_by(v - x0[d], sx) : v;
                }
                tile[tid * ST + D] = FAST ? scale_by(et[j0 + tid] - t0, st) : et[j0 + tid];
                if (FAST) tile[tid * ST + D + 1] = et[j0 + tid];
            }
            __syncthreads();
            if (wave_live) {
#pragma unroll
                for (int a = 0; a < A; ++a)
                    diag_pass<D, GRAD, F32, FAST, KT, PM>(tile, cnt, xi[a], ti[a], tn[a], c, acc[a]);
            }
        }
    } else {
        slot = chunk;
        const int c0 = chunk * ds.L;
        const int c1 = min(c0 + ds.L, P0);
        const bool tail = P0 + R > n;
        double *cp = colpart + ds.col0 + (size_t)k * ds.col_kstride + (size_t)NC * R * ((size_t)P * (P - 1) / 2);
        // flag of column tile t for this block: cf[t * nb]
        unsigned char *cf = L.colflag + ds.flag0 + (size_t)k * ds.flag_kstride + P;
        // the block's first row: the smallest time gap to any earlier tile is its gap
        const double t_blk = et[min(P0, n - 1)];
        // culled order: bounding boxes of the block's rows and of the row group this lane tests
        // (lane l < 4 A: the wave's row group l / 4; empty groups past n hold empty boxes)
        double blo[D + 1], bhi[D + 1], glo[D + 1], ghi[D + 1];
        const double hx = 0.5 * tauX * tauX, ht = 0.5 * tauT * tauT;
        if (PM == kPassCulledB) {
            group_box<D>(ds.gstat, P0 / kGroup, R / kGroup, blo, bhi);
            group_box<D>(ds.gstat, P0 / kGroup + row_group((lane >> 2) & (A - 1)), 1, glo, ghi);
        }
        bool wrote = false; // a tile of the item passed the workgroup-level skip tests (uniform)
        // Whether any wave can form an excitation term in the item: the per-wave test below skips
        // them for a tile when omega times the gap between the wave's first row and the tile's
        // last column passes the cut; the chunk's last tile and the block's first row (wave 0)
        // give the smallest gap, so when that pair passes the cut every wave skips every tile.
        // Otherwise the excitation moments are written, zeros included.
        full = PM != kPassCulledB;
        if (full && L.skip_zero) {
            const double tl = et[c1 - 1];
            if (tl < t_blk && gs >= 0.0 && gs * (t_blk - tl) >= zs) full = false;
        }
        const int box_tile0 = culled_groups(n), box_super0 = box_tile0 + culled_groups(n) * kGroup / kTile;
        // the rows' sums over kFlushTiles computed tiles are formed apart (tacc) and then added to
        // the item's (acc): a row's sum over a chunk of L columns grows its rounding as
        // sqrt(kFlushTiles kTile) + sqrt(L / (kFlushTiles kTile)) rather than sqrt(L) (gradient
        // components that cancel between events amplify it)
        double tacc[A][6]; // FLUSH only
#pragma unroll
        for (int a = 0; a < A; ++a)
#pragma unroll
            for (int q = 0; q < 6; ++q) tacc[a][q] = 0.0;
        int since = 0;
        for (int j0 = c0; j0 < c1; j0 += kTile) {
            // c1 - c0 and P0 are multiples of kTile: every tile here is full.
            if (PM == kPassCulledB && L.skip_zero) {
                // every background exponent between the block and a supertile (when one starts
                // here and lies inside the chunk) or the tile past the negligible point: their
                // column partials are zero
                double tlo[D + 1], thi[D + 1];
                if (j0 % kSuperTile == 0 && j0 + kSuperTile <= c1) {
                    group_box<D>(ds.gstat, box_super0 + j0 / kSuperTile, 1, tlo, thi);
                    if (box_background_exponent<D>(blo, bhi, tlo, thi, hx, ht) >= ds.zero_arg_b) {
                        if (tid < kSuperTile / kTile) cf[(size_t)(j0 / kTile + tid) * ds.nb] = 0;
                        j0 += kSuperTile - kTile;
                        continue;
                    }
                }
                group_box<D>(ds.gstat, box_tile0 + j0 / kTile, 1, tlo, thi);
                if (box_background_exponent<D>(blo, bhi, tlo, thi, hx, ht) >= ds.zero_arg_b) {
                    if (tid == 0) cf[(size_t)(j0 / kTile) * ds.nb] = 0;
                    continue;
                }
            }
            if (PM == kPassBoth && L.skip_zero) {
                // Workgroup-wide skip, decided as the per-wave test below decides it for the
                // first wave (whose gap is the smallest): when neither term survives there, it
                // survives for no row of the block; the tile's column partials are then zero
                // (flagged, not written).
                const double tl = et[j0 + kTile - 1];
                const double gap = t_blk - tl;
                if (tl < t_blk && gb * gap * gap >= zb && gs >= 0.0 && gs * gap >= zs) {
                    if (tid == 0) cf[(size_t)(j0 / kTile) * ds.nb] = 0;
                    continue;
                }
            }
            __syncthreads();
            // culled order: which of the tile's four 64-column groups lie past the negligible
            // point for which of the wave's row groups (bit 4 a + sb: row group a, subtile sb;
            // wave-uniform). Lane l < 4 A tests row group l / 4 against column group l mod 4; the
            // box loads are issued with the tile's and complete under the same barrier.
            int farmask = 0;
            static_assert(kGroup == 64, "a box group is one 64-column subtile");
            if (PM == kPassCulledB && L.skip_zero) {
                double slo[D + 1], shi[D + 1];
                group_box<D>(ds.gstat, j0 / kGroup + (lane & 3), 1, slo, shi);
                const bool far = box_background_exponent<D>(glo, ghi, slo, shi, hx, ht) >= ds.zero_arg_b;
                farmask = (int)(__ballot(far && lane < 4 * A) & ((1ULL << (4 * A)) - 1ULL));
            }
            {
                // subtile sb = tid / 64 holds columns j0 + 64 sb .. + 63 (SubStride); lanes then
                // read consecutive doubles.
                const int e0 = (tid >> 6) * SUB + (tid & 63);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double u = ex[(size_t)d * npad + j0 + tid];
                    const double v = FAST ? scale_by(u - x0[d], sx) : u;
                    tile[e0 + d * 128] = v;
                    tile[e0 + d * 128 + 64] = v;
                }
                const double tnat = et[j0 + tid];
                const double tv = FAST ? scale_by(tnat - t0, st) : tnat;
                tile[e0 + D * 128] = tv;
                tile[e0 + D * 128 + 64] = tv;
                if (FAST) tile[e0 + (D + 1) * 128] = tnat;
            }
            __syncthreads();
            if (wave_live) {
                // time-sorted passes: columns are earlier than the rows, tl <= t_lo; equal times
                // need the mask
                // the tile's last column, natural time (staged beside the scaled one in base 2)
                const double tl = tile[3 * SUB + (FAST ? D + 1 : D) * 128 + 63];
                int smode = PM == kPassCulledB ? kSNone : (tl < t_lo ? kSFull : kSMixed);
                int dob = 1;
                if (PM != kPassCulledB && L.skip_zero) {
                    const double gap = tl < t_lo ? t_lo - tl : 0.0;
                    if (gb * gap * gap >= zb) dob = 0;
                    if (smode == kSFull && gs >= 0.0 && gs * gap >= zs) smode = kSNone;
                }
                const int mode0 = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
                for (int sb = 0; sb < 4; ++sb) {
                    int mode = mode0;
                    const double *sub = tile + sb * SUB;
                    auto &ac = FLUSH ? tacc : acc;
                    if constexpr (PM == kPassCulledB && A == 2) {
                        // culled order: the subtile against the wave's row groups that are near it
                        // (both: one two-row call; one: a one-row call on that row's registers)
                        const bool near0 = !L.skip_zero || !((farmask >> sb) & 1);
                        const bool near1 = !L.skip_zero || !((farmask >> (4 + sb)) & 1);
                        if (near0 != near1) {
                            sym_one_row<D, GRAD, F32, FAST, KT, PM>(near1, tail, mode, sub, lane, xi, ti, tn, vm, c, ac, cs);
                            mode = 0;
                        } else if (!near0) {
                            mode = 0;
                        }
                    } else if (PM == kPassCulledB && L.skip_zero && ((farmask >> sb) & 1)) {
                        mode = 0; // one row group per wave: the subtile is far from it
                    }
                    if (mode != 0) {
                        if (tail)
                            sym_dispatch<D, GRAD, A, true, F32, FAST, KT, PM>(mode, sub, lane, xi, ti, tn, vm, c, ac, cs);
                        else
                            sym_dispatch<D, GRAD, A, false, F32, FAST, KT, PM>(mode, sub, lane, xi, ti, tn, vm, c, ac, cs);
                    }
                    // column lane of the subtile: its slots before and after the wrap; the slots
                    // are zeroed again for the next subtile (the wave's LDS operations take effect
                    // in program order)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        colbuf[(w * NC + c) * kTile + sb * 64 + lane] = cs[c * 128 + lane] + cs[c * 128 + lane + 64];
                        cs[c * 128 + lane] = 0.0;
                        cs[c * 128 + lane + 64] = 0.0;
                    }
                }
            } else {
#pragma unroll
                for (int sb = 0; sb < 4; ++sb)
#pragma unroll
                    for (int c = 0; c < NC; ++c) colbuf[(w * NC + c) * kTile + sb * 64 + lane] = 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double v = colbuf[(0 * NC + c) * kTile + tid];
#pragma unroll
                for (int ww = 1; ww < kSymWaves; ++ww) v += colbuf[(ww * NC + c) * kTile + tid];
                cp[(size_t)c * P0 + j0 + tid] = v;
            }
            if (tid == 0) cf[(size_t)(j0 / kTile) * ds.nb] = 1;
            wrote = true;
            if (FLUSH && ++since == kFlushTiles) {
                since = 0;
#pragma unroll
                for (int a = 0; a < A; ++a)
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        acc[a][q] += tacc[a][q];
                        tacc[a][q] = 0.0;
                    }
            }
        }
        if (FLUSH) {
#pragma unroll
            for (int a = 0; a < A; ++a)
#pragma unroll
                for (int q = 0; q < 6; ++q) acc[a][q] += tacc[a][q];
        }
        // the row-slot flag: an item whose every tile was skipped leaves its (zero) row partials
        // unwritten (0); one that could form no excitation term writes its background moments
        // only (1: the excitation moments are exactly zero); else all of them (2)
        unsigned char *rf = L.colflag + ds.rflag0 + (size_t)k * ds.rflag_kstride + (size_t)P * ds.max_chunks + chunk;
        if (tid == 0) *rf = wrote ? (full ? 2 : 1) : 0;
        if (!wrote) return;
    }
    const size_t rpad = (size_t)ds.nb * R;
    double *rp = rowpart + ds.row0 + (size_t)k * ds.row_kstride + (size_t)slot * NR * rpad;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int r = P0 + 64 * row_group(a) + lane;
        if (GRAD) {
#pragma unroll
            for (int q = 0; q < 3; ++q) rp[(size_t)q * rpad + r] = acc[a][q];
            if (full) {
#pragma unroll
                for (int q = 3; q < 6; ++q) rp[(size_t)q * rpad + r] = acc[a][q];
            }
        } else {
            rp[r] = acc[a][0];
            if (full) rp[rpad + r] = acc[a][3];
        }
    }
}

// Persistent pass: every workgroup takes work items from the queue until it is empty. FM = 0:
// every item in the natural form; 1: every item in the base-2 form; 2: each item in the base-2
// form where its dataset allows it (SymData::fast), else natural (both paths in one kernel, at
// some register cost, so it is used only for mixed dataset tables); 3: as 1, with every
// background exponent of the launch in exp2_neg_ktab's range (SymLaunch::ktab; FP64 only).
// LDS of sym_kernel in doubles (below): subtiles or the diagonal tile, column slots, the
// exponential's table (base-2 form in FP64), per-wave column sums.
template <int D, bool GRAD, bool F32, int FM> struct SymLds {
    static constexpr int SUB = SubStride<D, FM != 0>::value;
    static constexpr int ST = DiagStride<D, FM != 0>::value;
    static constexpr int NC = GRAD ? 3 : 1;
    static constexpr int tile = (4 * SUB > kTile * ST) ? 4 * SUB : kTile * ST;
    static constexpr int slot = kSymWaves * NC * 128;
    static constexpr int tab = FM != 0 ? kExp2TabSize : 0;
    static constexpr int total = tile + slot + tab + kSymWaves * NC * kTile;
    // workgroups of kSymWaves waves that fit one CU's 160 KB, at most 3 (waves per SIMD); 1 KB
    // per workgroup left for the compiler's own (the queue index, __syncthreads_or)
    static constexpr int waves = (3 * (8 * total + 1024) <= 160 * 1024) ? 3 : 2;
};

// Three waves per SIMD (at most 168 VGPRs): measured slightly faster than the two that the
// unconstrained allocation (up to 228 VGPRs for 2 rows per lane) leaves room for. Where the LDS
// allows two workgroups per CU only (D = 3 in the base-2 form), two waves per SIMD (256 VGPRs).
template <int D, bool GRAD, int A, bool F32, int FM, int PM>
__global__ __launch_bounds__(kSymWaves * 64) __attribute__((amdgpu_waves_per_eu(
    (FM == 2 || SymLds<D, GRAD, F32, FM>::waves == 2) ? 2 : 3, SymLds<D, GRAD, F32, FM>::waves))) void sym_kernel(
    SymLaunch L, const double *__restrict__ params, double *__restrict__ rowpart, double *__restrict__ colpart) {
    using Lds = SymLds<D, GRAD, F32, FM>;
    constexpr int NC = GRAD ? 3 : 1;
    constexpr bool TAB = FM != 0; // the base-2 form: exp2_neg_tab (FP64 background, excitation)
    constexpr bool KT = FM == 3;          // exp2_neg_ktab and its table
    static_assert(!KT || !F32, "the exponent-in-table form is FP64's");
    // diagonal items: kTile events x ST doubles; chunk items: 4 subtiles x SubStride doubles
    constexpr int kTileD = Lds::tile;
    constexpr int kSlotD = Lds::slot;
    constexpr int kTabD = Lds::tab;
    // One LDS array, laid out by hand: the arrays the pair loop reads (subtiles, column slots,
    // the exponential's table) first, so that their addresses fit the 16-bit offset field of the
    // LDS instructions (no address additions in the loop); the per-wave column sums last.
    __shared__ double lds[Lds::total];
    double *const tile = lds;
    double *const colslot = lds + kTileD;
    const double *const etab = lds + kTileD + kSlotD;
    double *const colbuf = lds + kTileD + kSlotD + kTabD;
    __shared__ int s_item;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // the table is complete before the first item: the queue loop's barrier follows
    if (TAB) fill_exp2_table(lds + kTileD + kSlotD, tid, kSymWaves * 64, KT);
    // this wave's column slots (sym_subtile), zero between subtiles
    double *cs = colslot + w * NC * 128;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        cs[c * 128 + lane] = 0.0;
        cs[c * 128 + lane + 64] = 0.0;
    }
    for (;;) {
        const int idx = queue_next(L, &s_item);
        if (idx >= L.items) break;
        const int4 it = L.item[idx];
        const bool fast = FM == 1 || FM == 3 || (FM == 2 && L.data[it.x].fast);
        const bool flush = it.z >= 0 && L.data[it.x].L > kFlushTiles * kTile;
        if (fast) {
            if (flush)
                sym_item<D, GRAD, A, F32, true, KT, PM, true>(L, it, params, rowpart, colpart, tile, colbuf, cs, etab, tid, lane, w);
            else
                sym_item<D, GRAD, A, F32, true, KT, PM, false>(L, it, params, rowpart, colpart, tile, colbuf, cs, etab, tid, lane, w);
        } else if constexpr (FM != 1 && FM != 3) {
            if (flush)
                sym_item<D, GRAD, A, F32, false, false, PM, true>(L, it, params, rowpart, colpart, tile, colbuf, cs, etab, tid, lane, w);
            else
                sym_item<D, GRAD, A, F32, false, false, PM, false>(L, it, params, rowpart, colpart, tile, colbuf, cs, etab, tid, lane, w);
        }
    }
}

template <int D, bool GRAD, bool F32, int FM, int PM>
static void launch_sym_d(const SymLaunch &L, const double *params, double *rowpart, double *colpart, hipStream_t s) {
    if (L.A == 2)
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 2, F32, FM, PM>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s,
                           L, params, rowpart, colpart);
    else
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 1, F32, FM, PM>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s,
                           L, params, rowpart, colpart);
}

template <bool F32, int FM, int PM>
static void launch_sym_p(const SymLaunch &L, const double *params, bool grad, double *rowpart, double *colpart,
                         hipStream_t s) {
    switch (L.dim * 2 + (grad ? 1 : 0)) {
    case 2: launch_sym_d<1, false, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    case 3: launch_sym_d<1, true, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    case 4: launch_sym_d<2, false, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    case 5: launch_sym_d<2, true, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    case 6: launch_sym_d<3, false, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    case 7: launch_sym_d<3, true, F32, FM, PM>(L, params, rowpart, colpart, s); break;
    default: throw std::invalid_argument("symmetric pass is built for dimensions 1..3");
    }
}

// The culled background pass (pass mode kPassCulledB; FM 0 or 1), built in hph_split.hip.
void launch_sym_split(const SymLaunch &L, const double *params, bool grad, bool fp32, double *rowpart,
                      double *colpart, hipStream_t s);

} // namespace kern
} // namespace hph
