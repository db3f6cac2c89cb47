// HIP kernels of the Hawkes likelihood engine, written for gfx950 (CDNA4, wave64).
//
// The O(N^2) pass restates the per-row rate sum of src/NewHawkes.hpp:520-547 (ratesLoop) and
// src/backend/opencl/OpenCLHawkes.hpp:632-702 (computeLikContribs), with three changes of
// method:
//   * one event row per lane, the column events staged in LDS tiles and read by broadcast;
//   * the two Gaussians of the background term share one exp, and so do the Gaussian and the
//     exponential of the excitation term: pdf(a)pdf(b) = exp(-(a^2+b^2)/2)/(2 pi);
//   * events are time-sorted, so the excitation term (dt > 0 only) is skipped for whole tiles
//     that lie after a wavefront's rows and needs no mask for tiles wholly before them.
// Per-row terms (log, exp and Phi integral terms, src/NewHawkes.hpp:652-655) and the exact
// gradient are formed in a second kernel; all sums over rows are fixed-order double-double
// reductions, so results are bitwise reproducible run to run.

#include "hph_device.h"

#include <stdexcept>

namespace hph {
namespace kern {


// One LDS tile against this lane's row. DOB says whether the background term is formed (it is
// not when every background term of the tile is exactly zero). SMODE says whether the
// excitation term is absent (every column at or after the row block, or every term exactly
// zero), unmasked (every column strictly before it) or masked per pair (the diagonal band, and
// ties).
template <int D, bool GRAD, bool DOB, int SMODE, bool F32>
__device__ __forceinline__ void tile_pass(const double *__restrict__ tile, int cnt, const double (&xi)[D],
                                          double ti, double cBx, double cBt, double cSx, double nomega,
                                          double (&acc)[6]) {
    constexpr int ST = TileStride<D>::value;
#pragma unroll 2
    for (int j = 0; j < cnt; ++j) {
        const double *e = tile + j * ST;
        double r2;
        {
            const double d0 = xi[0] - e[0];
            r2 = d0 * d0;
        }
#pragma unroll
        for (int d = 1; d < D; ++d) {
            const double dd = xi[d] - e[d];
            r2 = __builtin_fma(dd, dd, r2);
        }
        const double dt = ti - e[D];
        if (DOB) {
            const double dt2 = dt * dt;
            const double eb = background_exp<F32>(cBx, cBt, r2, dt2);
            acc[0] += eb;
            if (GRAD) {
                acc[1] = __builtin_fma(eb, r2, acc[1]);
                acc[2] = __builtin_fma(eb, dt2, acc[2]);
            }
        }
        if (SMODE != kSNone) {
            double es = exp_fp64(__builtin_fma(cSx, r2, nomega * dt));
            if (SMODE == kSMixed) es = dt > 0.0 ? es : 0.0;
            acc[3] += es;
            if (GRAD) {
                acc[4] = __builtin_fma(es, r2, acc[4]);
                acc[5] = __builtin_fma(es, dt, acc[5]);
            }
        }
    }
}

// Grid: x = row_blocks * chunks (later row blocks first: they carry more excitation work),
// y = parameter vector k. Block: 256 lanes, one row each.
template <int D, bool GRAD, bool F32>
__global__ __launch_bounds__(kRowsPerBlock) void pair_kernel(Events ev, const double *__restrict__ params,
                                                             PairGrid g, double *__restrict__ part) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NACC = GRAD ? 6 : 2;
    __shared__ double tile[kTile * ST];

    const int k = blockIdx.y;
    const int rb = g.row_blocks - 1 - (int)(blockIdx.x / g.chunks);
    const int c = (int)(blockIdx.x % g.chunks);
    const double *pk = params + 6 * k;
    const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3];
    const double cBx = -0.5 * tauX * tauX, cBt = -0.5 * tauT * tauT;
    const double cSx = -0.5 * sigma * sigma, nomega = -omega;
    // background-exponential constants (log2 e scaled for the FP32-bulk v_exp_f32)
    const double eBx = F32 ? cBx * kLog2e : cBx, eBt = F32 ? cBt * kLog2e : cBt;
    const double zs = g.cut_s ? excitation_skip_arg(pk, D, ev.n) : g.zero_arg_s;

    const int tid = threadIdx.x;
    const int n = ev.n, npad = ev.npad;
    const int i = rb * kRowsPerBlock + tid;
    const int ic = min(i, n - 1);
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = ev.x[(size_t)d * npad + ic];
    const double ti = ev.t[ic];

    const int wave_first = __builtin_amdgcn_readfirstlane(rb * kRowsPerBlock + (tid & ~63));
    const bool wave_live = wave_first < n;
    const double t_lo = ev.t[min(wave_first, n - 1)];
    const double t_hi = ev.t[min(wave_first + 63, n - 1)];

    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int j_begin = c * g.chunk_len;
    const int j_end = min(n, j_begin + g.chunk_len);
    for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
        const int cnt = min(kTile, j_end - j0);
        __syncthreads();
        if (tid < cnt) {
#pragma unroll
            for (int d = 0; d < D; ++d) tile[tid * ST + d] = ev.x[(size_t)d * npad + j0 + tid];
            tile[tid * ST + D] = ev.t[j0 + tid];
        }
        __syncthreads();
        if (wave_live) {
            const double tf = tile[D], tl = tile[(cnt - 1) * ST + D];
            int smode = tl < t_lo ? kSFull : (tf >= t_hi ? kSNone : kSMixed);
            int dob = 1;
            if (g.skip_zero) {
                // Smallest |dt| between this wavefront's rows and the tile (times are sorted).
                const double gap = tl < t_lo ? t_lo - tl : (tf > t_hi ? tf - t_hi : 0.0);
                // Background: arg <= -tauT^2 dt^2 / 2, skipped below the negligible point
                // (background_skip_arg); excitation (tile before the rows only): arg <= -omega dt,
                // skipped where every term is exactly +0.0.
                if (-cBt * gap * gap >= g.zero_arg_b) dob = 0;
                if (smode == kSFull && nomega <= 0.0 && -nomega * gap >= zs) smode = kSNone;
            }
            const int mode = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
            switch (mode) {
            case 4 + kSNone: tile_pass<D, GRAD, true, kSNone, F32>(tile, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
            case 4 + kSFull: tile_pass<D, GRAD, true, kSFull, F32>(tile, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
            case 4 + kSMixed: tile_pass<D, GRAD, true, kSMixed, F32>(tile, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
            case kSFull: tile_pass<D, GRAD, false, kSFull, F32>(tile, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
            case kSMixed: tile_pass<D, GRAD, false, kSMixed, F32>(tile, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
            default: break; // every term of the tile is exactly zero
            }
        }
    }
    double *out = part + ((size_t)k * g.chunks + c) * NACC * npad;
    if (GRAD) {
#pragma unroll
        for (int a = 0; a < 6; ++a) out[(size_t)a * npad + i] = acc[a];
    } else {
        out[i] = acc[0];
        out[npad + i] = acc[3];
    }
}

// Narrow layout: one workgroup per (64-row group, 256-column chunk, vector), grid (chunks, row
// groups, K). The four wavefronts share the 64 rows and take 64 columns each, so the per-lane
// loop is 64 pairs long; their sums are added in LDS in a fixed order into one partial per
// (chunk, row), the partial layout of pair_kernel.
template <int D, bool GRAD, bool F32>
__global__ __launch_bounds__(kRowsPerBlock) void pair_kernel_narrow(Events ev, const double *__restrict__ params,
                                                                    PairGrid g, double *__restrict__ part) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NACC = GRAD ? 6 : 2;
    __shared__ double tile[kTile * ST];
    __shared__ double red[3][NACC][kNarrow];

    const int c = blockIdx.x, rg = blockIdx.y, k = blockIdx.z;
    const double *pk = params + 6 * k;
    const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3];
    const double cBx = -0.5 * tauX * tauX, cBt = -0.5 * tauT * tauT;
    const double cSx = -0.5 * sigma * sigma, nomega = -omega;
    const double eBx = F32 ? cBx * kLog2e : cBx, eBt = F32 ? cBt * kLog2e : cBt;
    const double zs = g.cut_s ? excitation_skip_arg(pk, D, ev.n) : g.zero_arg_s;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = ev.n, npad = ev.npad;
    const int i = rg * kNarrow + lane;
    const int ic = min(i, n - 1);
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = ev.x[(size_t)d * npad + ic];
    const double ti = ev.t[ic];
    const int r0 = rg * kNarrow;
    const double t_lo = ev.t[r0];
    const double t_hi = ev.t[min(r0 + kNarrow - 1, n - 1)];

    const int j0 = c * kTile;
    const int cnt_all = min(kTile, n - j0);
    if (tid < cnt_all) {
#pragma unroll
        for (int d = 0; d < D; ++d) tile[tid * ST + d] = ev.x[(size_t)d * npad + j0 + tid];
        tile[tid * ST + D] = ev.t[j0 + tid];
    }
    __syncthreads();
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int cnt = min(kNarrow, cnt_all - w * kNarrow); // this wavefront's columns (may be <= 0)
    if (cnt > 0) {
        const double *sub = tile + w * kNarrow * ST;
        const double tf = sub[D], tl = sub[(cnt - 1) * ST + D];
        int smode = tl < t_lo ? kSFull : (tf >= t_hi ? kSNone : kSMixed);
        int dob = 1;
        if (g.skip_zero) {
            const double gap = tl < t_lo ? t_lo - tl : (tf > t_hi ? tf - t_hi : 0.0);
            if (-cBt * gap * gap >= g.zero_arg_b) dob = 0;
            if (smode == kSFull && nomega <= 0.0 && -nomega * gap >= zs) smode = kSNone;
        }
        switch (dob * 4 + smode) {
        case 4 + kSNone: tile_pass<D, GRAD, true, kSNone, F32>(sub, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
        case 4 + kSFull: tile_pass<D, GRAD, true, kSFull, F32>(sub, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
        case 4 + kSMixed: tile_pass<D, GRAD, true, kSMixed, F32>(sub, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
        case kSFull: tile_pass<D, GRAD, false, kSFull, F32>(sub, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
        case kSMixed: tile_pass<D, GRAD, false, kSMixed, F32>(sub, cnt, xi, ti, eBx, eBt, cSx, nomega, acc); break;
        default: break; // every term of the sub-tile is skipped
        }
    }
    constexpr int IDX[6] = {0, GRAD ? 1 : 3, 2, 3, 4, 5}; // accumulators kept: all 6, or sums B and S
    if (w > 0) {
#pragma unroll
        for (int a = 0; a < NACC; ++a) red[w - 1][a][lane] = acc[IDX[a]];
    }
    __syncthreads();
    if (w == 0) {
        double *out = part + ((size_t)k * g.chunks + c) * NACC * npad;
#pragma unroll
        for (int a = 0; a < NACC; ++a) {
            double v = acc[IDX[a]];
#pragma unroll
            for (int q = 0; q < 3; ++q) v += red[q][a][lane];
            out[(size_t)a * npad + i] = v;
        }
    }
}

// Per-row terms: ll_i = log(lambda_i) + theta (e^{-omega (T - t_i)} - 1)
//                       - mu0 (Phi(tauT (T - t_i)) - Phi(-tauT t_i))   (src/NewHawkes.hpp:652-655)
// and d ll_i / d(sigma, tauX, tauT, omega, theta, mu0).
template <bool GRAD>
__global__ __launch_bounds__(kRowsPerBlock) void finalize_kernel(Events ev, const double *__restrict__ params,
                                                                 PairGrid g, const double *__restrict__ part,
                                                                 double *__restrict__ block_out,
                                                                 double *__restrict__ probs,
                                                                 double *__restrict__ inv_lam,
                                                                 double *__restrict__ sums) {
    constexpr int NACC = GRAD ? 6 : 2;
    __shared__ double2 sh[kResultWidth * kRowsPerBlock];
    const int k = blockIdx.y;
    const int rb = blockIdx.x;
    const int i = rb * kRowsPerBlock + threadIdx.x;
    const int n = ev.n, npad = ev.npad;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        double a6[NACC];
#pragma unroll
        for (int a = 0; a < NACC; ++a) a6[a] = 0.0;
#pragma unroll 4
        for (int c = 0; c < g.chunks; ++c) { // unrolled: several chunks' loads in flight
            const double *base = part + ((size_t)k * g.chunks + c) * NACC * npad;
#pragma unroll
            for (int a = 0; a < NACC; ++a) a6[a] += base[(size_t)a * npad + i];
        }
        double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (GRAD) {
#pragma unroll
            for (int a = 0; a < 6; ++a) s[a] = a6[a];
        } else {
            s[0] = a6[0];
            s[3] = a6[NACC - 1];
        }
        if (sums != nullptr && k == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (GRAD || q == 0 || q == 3) sums[(size_t)q * npad + i] = s[q];
        }
        double lam;
        const double T = ev.t_last_at != nullptr ? *ev.t_last_at : ev.t_last;
        const double p = row_terms<GRAD>(s, params + 6 * k, ev.dim, ev.t[i], T, v, lam);
        if (probs != nullptr && k == 0) probs[ev.perm[i]] = p;
        if (inv_lam != nullptr && k == 0) inv_lam[i] = 1.0 / lam;
    }
    store_block_sums<GRAD>(v, sh, block_out, k, g.row_blocks, rb);
}

// One block per k: fixed-order double-double sum of the block partials, plus the
// N (D-1) log(1/sqrt(2 pi)) constant (src/NewHawkes.hpp:659).
template <int W>
__global__ __launch_bounds__(kRowsPerBlock) void reduce_kernel(Events ev, int nblocks,
                                                               const double *__restrict__ block_out,
                                                               double *__restrict__ result) {
    __shared__ double2 sh[W * kRowsPerBlock];
    const int k = blockIdx.x;
    double2 a[W];
#pragma unroll
    for (int q = 0; q < W; ++q) a[q] = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nblocks; b += kRowsPerBlock) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const double *o = block_out + (((size_t)k * nblocks + b) * kResultWidth + q) * 2;
            a[q] = dd_add(a[q], make_double2(o[0], o[1]));
        }
    }
    block_reduce_dd_w<W>(a, sh);
    if (threadIdx.x == 0) {
        const double c = -0.91893853320467274178 * (double)ev.n * (double)(ev.dim - 1);
        a[0] = dd_add(a[0], make_double2(c, 0.0));
#pragma unroll
        for (int q = 0; q < W; ++q) result[(size_t)k * kResultWidth + q] = a[q].x + a[q].y;
    }
}

// Excitation sums per event alone, for re-evaluations whose background sums are cached: one
// workgroup per 64 events (one per lane); the earlier events' 256-event tiles are visited from the
// events' own tile backwards, each split over the four wavefronts (64 columns each), until every
// remaining excitation term is past the likelihood passes' cut (excitation_skip_arg: omega times
// the smallest time gap at or past X_S). Natural form, FP64, per-pair dt > 0 mask (ties and later
// events contribute nothing); the wavefronts' sums are added in a fixed order. out: [3 or 1][npad]
// in natural units (S, S r^2, S dt), per vector k (grid y): out + k NQ npad. cut: 0 the
// likelihood passes' point X_S, 1 exactly zero terms only (omega gap >= 746, probability
// requests), 2 none (every earlier event). For omega >= 0 (every causal exponent
// -(sigma^2 r^2 / 2 + omega dt) <= 0) the terms come from the LDS table exponential exp2_neg_tab,
// the exponent scaled by 256 log2(e) with double-double constants (the scale of sigma^2 / 2 and
// omega the finalize's): 10 VALU operations against exp_fp64's 16; omega < 0 keeps exp_fp64.
template <int D, bool GRAD>
__global__ __launch_bounds__(kRowsPerBlock) void excitation_band_kernel(Events ev, const double *__restrict__ params,
                                                                        int cut, double *__restrict__ out) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NQ = GRAD ? 3 : 1;
    __shared__ double tile[kTile * ST];
    __shared__ double red[3][NQ][kNarrow];
    const int rg = blockIdx.x, k = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = ev.n, npad = ev.npad;
    const int i = rg * kNarrow + lane;
    const int ic = min(i, n - 1);
    const double *pk = params + 6 * (size_t)k;
    const double sigma = pk[0], omega = pk[3];
    const double cSx = -0.5 * sigma * sigma;
    const double xs = cut == 0 ? excitation_skip_arg(pk, D, n) : (cut == 1 || cut == 3 ? kZeroExpArg : HUGE_VAL);
    // cut 3 (sums without moments): each row also stops, relative to its own sum, once every
    // earlier event left could add no more than 2^-60 of it (below)
    double sD = 1.0, xD = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        sD *= sigma;
        xD *= pk[1];
    }
    // only where the rate is a sum of non-negative parts, so that a bound relative to the
    // excitation sum is one relative to the rate and to the probability
    const bool rel = !GRAD && cut == 3 && pk[4] * omega * sD > 0.0 && pk[5] * xD * pk[2] >= 0.0;
    __shared__ double part[4][kNarrow];
    __shared__ double etab[kExp2TabSize];
    fill_exp2_table(etab, tid, kRowsPerBlock); // complete before the first tile's barriers
    const bool tabf = omega >= 0.0;
    // yS = c0 r^2 + c1 dt = 256 log2(e) (sigma^2 r^2 / 2 + omega dt), constants as double-doubles
    const double kh = 256.0 * kLog2eHi, kl = 256.0 * kLog2eLo;
    const double s2h = -cSx;
    const double c0 = s2h * kh, c0lo = __builtin_fma(s2h, kh, -c0) + s2h * kl;
    const double c1 = omega * kh, c1lo = __builtin_fma(omega, kh, -c1) + omega * kl;
    bool done = false;
    out += (size_t)k * NQ * npad;
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = ev.x[(size_t)d * npad + ic];
    const double ti = ev.t[ic];
    const double t_first = ev.t[rg * kNarrow]; // the group's smallest time (sorted)
    double acc[3] = {0.0, 0.0, 0.0};
    const int t_own = rg * kNarrow / kTile;    // the tile holding the group's rows
    for (int tj = t_own; tj >= 0; --tj) {
        const int j0 = tj * kTile;
        if (tj < t_own && omega >= 0.0 && omega * (t_first - ev.t[j0 + kTile - 1]) >= xs) break; // uniform
        if (done) break; // uniform: every wave took the same vote
        // columns up to the group's last row: later events excite none of its rows
        const int cnt_all = min(min(kTile, n - j0), (rg + 1) * kNarrow - j0);
        __syncthreads();
        if (tid < cnt_all) {
#pragma unroll
            for (int d = 0; d < D; ++d) tile[tid * ST + d] = ev.x[(size_t)d * npad + j0 + tid];
            tile[tid * ST + D] = ev.t[j0 + tid];
        }
        __syncthreads();
        const int cnt = min(kNarrow, cnt_all - w * kNarrow); // this wavefront's columns (may be <= 0)
        const double *sub = tile + w * kNarrow * ST;
#pragma unroll 2
        for (int j = 0; j < cnt; ++j) {
            const double *e = sub + j * ST;
            double r2;
            {
                const double d0 = xi[0] - e[0];
                r2 = d0 * d0;
            }
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const double dd = xi[d] - e[d];
                r2 = __builtin_fma(dd, dd, r2);
            }
            const double dt = ti - e[D];
            double es;
            if (tabf) // (uniform) a pair with dt <= 0 may give +inf or a wrong value: masked below
                es = exp2_neg_tab(__builtin_fma(c0, r2, __builtin_fma(c0lo, r2, __builtin_fma(c1, dt, c1lo * dt))), etab);
            else
                es = exp_fp64(__builtin_fma(cSx, r2, -omega * dt));
            es = dt > 0.0 ? es : 0.0;
            acc[0] += es;
            if (GRAD) {
                acc[1] = __builtin_fma(es, r2, acc[1]);
                acc[2] = __builtin_fma(es, dt, acc[2]);
            }
        }
        if (rel && j0 > 0) {
            // The j0 events still ahead are no later than t[j0 - 1]: each of their terms is at most
            // exp(-omega (t_i - t[j0 - 1])) (the spatial factor is at most 1), so together they
            // change row i's sum by at most j0 exp(-omega gap) <= 2^-60 of the sum formed so far
            // once the test below holds: a relative bound for every row, tiny sums included. The
            // rows' sums so far are the four waves' partials, added in a fixed order; every wave
            // evaluates the same test for the same 64 rows, so the vote is uniform.
            part[w][lane] = acc[0];
            __syncthreads();
            const double sum = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
            const double gap = ti - ev.t[j0 - 1];
            const bool ok = i >= n || (omega > 0.0 && (double)j0 * exp_fp64(-omega * gap) <= 0x1p-60 * sum);
            done = __all(ok) != 0;
        }
    }
    if (w > 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) red[w - 1][q][lane] = acc[q];
    }
    __syncthreads();
    if (w == 0 && i < n) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double v = acc[q];
#pragma unroll
            for (int u = 0; u < 3; ++u) v += red[u][q][lane];
            out[(size_t)q * npad + i] = v;
        }
    }
}

// ---------------- host launchers ----------------

PairGrid plan_pair_grid(int n, int K, int target_blocks) {
    PairGrid g;
    g.row_blocks = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    const int max_chunks = (n + kTile - 1) / kTile;
    g.skip_zero = 1;
    g.zero_arg_b = background_skip_arg(n, false);
    g.zero_arg_s = kZeroExpArg;
    g.cut_s = 0;
    // Narrow layout when the widest split of 256 x 256 items reaches under a quarter of the
    // requested grid: 64-row groups x 256-column chunks (the widest split's partials).
    const long long full = (long long)K * g.row_blocks * max_chunks;
    if (4 * full < target_blocks) {
        g.narrow = 1;
        g.chunk_len = kTile;
        g.chunks = max_chunks;
        return g;
    }
    g.narrow = 0;
    long long per_k = (long long)K * g.row_blocks;
    int want = (int)((target_blocks + per_k - 1) / per_k);
    if (want < 1) want = 1;
    int chunks = want < max_chunks ? want : max_chunks;
    const int per_chunk = (n + chunks - 1) / chunks;
    g.chunk_len = ((per_chunk + kTile - 1) / kTile) * kTile;
    g.chunks = (n + g.chunk_len - 1) / g.chunk_len;
    return g;
}

template <int D, bool F32>
static void launch_pair_d(const Events &ev, const double *params, int K, const PairGrid &g, bool grad,
                          double *part, hipStream_t s) {
    if (g.narrow) {
        dim3 grid((unsigned)g.chunks, (unsigned)((ev.n + kNarrow - 1) / kNarrow), (unsigned)K);
        if (grad)
            hipLaunchKernelGGL((pair_kernel_narrow<D, true, F32>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g,
                               part);
        else
            hipLaunchKernelGGL((pair_kernel_narrow<D, false, F32>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g,
                               part);
        return;
    }
    dim3 grid((unsigned)(g.row_blocks * g.chunks), (unsigned)K);
    if (grad)
        hipLaunchKernelGGL((pair_kernel<D, true, F32>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, part);
    else
        hipLaunchKernelGGL((pair_kernel<D, false, F32>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, part);
}

template <bool F32>
static void launch_pair_p(const Events &ev, const double *params, int K, const PairGrid &g, bool grad,
                          double *part, hipStream_t s) {
    switch (ev.dim) {
    case 1: launch_pair_d<1, F32>(ev, params, K, g, grad, part, s); break;
    case 2: launch_pair_d<2, F32>(ev, params, K, g, grad, part, s); break;
    case 3: launch_pair_d<3, F32>(ev, params, K, g, grad, part, s); break;
    case 4: launch_pair_d<4, F32>(ev, params, K, g, grad, part, s); break;
    case 5: launch_pair_d<5, F32>(ev, params, K, g, grad, part, s); break;
    case 6: launch_pair_d<6, F32>(ev, params, K, g, grad, part, s); break;
    case 7: launch_pair_d<7, F32>(ev, params, K, g, grad, part, s); break;
    case 8: launch_pair_d<8, F32>(ev, params, K, g, grad, part, s); break;
    default: throw std::invalid_argument("embedding dimension must be 1..8");
    }
}

void launch_pair(const Events &ev, const double *params, int K, const PairGrid &g, bool grad, bool fp32,
                 double *part, hipStream_t s) {
    if (fp32)
        launch_pair_p<true>(ev, params, K, g, grad, part, s);
    else
        launch_pair_p<false>(ev, params, K, g, grad, part, s);
}

void launch_excitation_band(const Events &ev, const double *params, int K, bool grad, int cut, double *out,
                            hipStream_t s) {
    if (K < 1 || K > 65535) throw std::invalid_argument("excitation pass: 1..65535 vectors");
    if (cut < 0 || cut > 3) throw std::invalid_argument("excitation pass: cut mode 0, 1, 2 or 3");
    const dim3 grid((unsigned)((ev.n + kNarrow - 1) / kNarrow), (unsigned)K);
    const int ex = cut;
    switch (ev.dim * 2 + (grad ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((excitation_band_kernel<1, false>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    case 3: hipLaunchKernelGGL((excitation_band_kernel<1, true>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    case 4: hipLaunchKernelGGL((excitation_band_kernel<2, false>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    case 5: hipLaunchKernelGGL((excitation_band_kernel<2, true>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    case 6: hipLaunchKernelGGL((excitation_band_kernel<3, false>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    case 7: hipLaunchKernelGGL((excitation_band_kernel<3, true>), grid, dim3(kRowsPerBlock), 0, s, ev, params, ex, out); break;
    default: throw std::invalid_argument("the excitation-only pass is built for dimensions 1..3");
    }
}

void launch_finalize(const Events &ev, const double *params, int K, const PairGrid &g, bool grad,
                     const double *part, double *block_out, double *probs, double *inv_lam, double *sums,
                     hipStream_t s) {
    dim3 grid((unsigned)g.row_blocks, (unsigned)K);
    if (grad)
        hipLaunchKernelGGL((finalize_kernel<true>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, part,
                           block_out, probs, inv_lam, sums);
    else
        hipLaunchKernelGGL((finalize_kernel<false>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, part,
                           block_out, probs, inv_lam, sums);
}

void launch_reduce(const Events &ev, int K, int nblocks, bool grad, const double *block_out, double *result,
                   hipStream_t s) {
    if (grad)
        hipLaunchKernelGGL((reduce_kernel<kResultWidth>), dim3((unsigned)K), dim3(kRowsPerBlock), 0, s, ev, nblocks,
                           block_out, result);
    else
        hipLaunchKernelGGL((reduce_kernel<1>), dim3((unsigned)K), dim3(kRowsPerBlock), 0, s, ev, nblocks, block_out,
                           result);
}

} // namespace kern
} // namespace hph
