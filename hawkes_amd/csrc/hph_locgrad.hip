// Location gradient d ll / d x (N x D) of the Hawkes log-likelihood, written for gfx950 (wave64).
//
// The reference declares a location gradient for its Java callers
// (src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96) and never implements it (the .cpp
// leaves it commented out, :80-89). This is the exact derivative of the model of
// src/NewHawkes.hpp:520-547, 652-659. The integral terms do not depend on locations; with
// w_i = 1 / lambda_i, the exp forms eB = exp(-(tauX^2 r^2 + tauT^2 dt^2)/2) and
// eS = exp(-omega dt - sigma^2 r^2/2) (dt > 0), lambda_i = cB sum_j eB_ij + cS sum_j eS_ij with
// cB = mu0 tauX^D tauT / (2 pi), cS = theta omega sigma^D / sqrt(2 pi), and
// d eB_ij / d x_i = -tauX^2 (x_i - x_j) eB_ij, d eS_ij / d x_i = -sigma^2 (x_i - x_j) eS_ij:
//   d ll / d x_k = - sum_j (x_k - x_j) [cB tauX^2 eB_kj (w_k + w_j) + cS sigma^2 (eS_kj w_k + eS_jk w_j)].
//
// Two passes: the row-wise rate pass (pair_kernel without gradient) gives sum eB and sum eS per
// event, inv_rate_kernel turns them into w; locgrad_kernel then runs over ordered pairs (one
// event row per lane, column tiles with their w in LDS) and locgrad_finalize_kernel sums the
// chunk partials in a fixed order and scatters them to the caller's event order. Tiles are
// skipped only where every term is exactly +0.0 (the gradient of an event has no floor that
// would bound a negligible-term cut).
#include "hph_device.h"

#include <stdexcept>

namespace hph {
namespace kern {

// Excitation direction of a tile relative to a wavefront's rows.
enum { kLNone = 0, kLBefore = 1, kLAfter = 2, kLMixed = 3 };

template <int D, bool DOB, int SMODE>
__device__ __forceinline__ void locgrad_tile(const double *__restrict__ tile, int cnt, const double (&xk)[D],
                                             double tk, double wk, double cBx, double cBt, double cSx, double omega,
                                             double fB, double fS, double (&acc)[D]) {
    constexpr int LS = LocStride<D>::value;
#pragma unroll 2
    for (int j = 0; j < cnt; ++j) {
        const double *e = tile + j * LS;
        double dx[D];
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            dx[d] = xk[d] - e[d];
            r2 = __builtin_fma(dx[d], dx[d], r2);
        }
        const double dt = tk - e[D];
        const double wj = e[D + 1];
        double f = 0.0;
        if (DOB) f = fB * background_exp<false>(cBx, cBt, r2, dt * dt) * (wk + wj);
        if (SMODE != kLNone) {
            // the earlier event of the pair excites the later one: eS_kj (dt > 0, weight w_k) or
            // eS_jk (dt < 0, weight w_j); equal times excite neither
            const double es = exp_fp64(__builtin_fma(cSx, r2, -omega * fabs(dt)));
            double ws;
            if (SMODE == kLBefore) ws = wk;
            else if (SMODE == kLAfter) ws = wj;
            else ws = dt > 0.0 ? wk : (dt < 0.0 ? wj : 0.0);
            f = __builtin_fma(fS * es, ws, f);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = __builtin_fma(-f, dx[d], acc[d]);
    }
}

// w_i = 1 / lambda_i (sorted order) from the rate pass's chunk partials [chunks][2][npad].
__global__ __launch_bounds__(kRowsPerBlock) void inv_rate_kernel(Events ev, const double *__restrict__ params,
                                                                 PairGrid g, const double *__restrict__ part,
                                                                 double *__restrict__ w) {
    const int i = blockIdx.x * kRowsPerBlock + threadIdx.x;
    const int n = ev.n, npad = ev.npad;
    if (i >= npad) return;
    double sb = 0.0, ss = 0.0;
    if (i < n) {
#pragma unroll 4
        for (int c = 0; c < g.chunks; ++c) {
            const double *base = part + (size_t)c * 2 * npad;
            sb += base[i];
            ss += base[npad + i];
        }
    }
    const double sigma = params[0], tauX = params[1], tauT = params[2], omega = params[3], theta = params[4],
                 mu0 = params[5];
    const double cB = mu0 * ipow(tauX, ev.dim) * tauT * 0.15915494309189533577;
    const double cS = theta * omega * ipow(sigma, ev.dim) * 0.39894228040143267794;
    w[i] = i < n ? 1.0 / (cB * sb + cS * ss) : 0.0;
}

// Grid: x = row_blocks * chunks. Block: 256 lanes, one row each. part: [chunks][D][npad].
template <int D>
__global__ __launch_bounds__(kRowsPerBlock) void locgrad_kernel(Events ev, const double *__restrict__ params,
                                                                const double *__restrict__ w, PairGrid g,
                                                                double *__restrict__ part) {
    constexpr int LS = LocStride<D>::value;
    __shared__ double tile[kTile * LS];

    const int rb = g.row_blocks - 1 - (int)(blockIdx.x / g.chunks);
    const int c = (int)(blockIdx.x % g.chunks);
    const double sigma = params[0], tauX = params[1], tauT = params[2], omega = params[3], theta = params[4],
                 mu0 = params[5];
    const double cBx = -0.5 * tauX * tauX, cBt = -0.5 * tauT * tauT, cSx = -0.5 * sigma * sigma;
    const double fB = mu0 * ipow(tauX, D) * tauT * 0.15915494309189533577 * tauX * tauX;
    const double fS = theta * omega * ipow(sigma, D) * 0.39894228040143267794 * sigma * sigma;

    const int tid = threadIdx.x;
    const int n = ev.n, npad = ev.npad;
    const int i = rb * kRowsPerBlock + tid;
    const int ic = min(i, n - 1);
    double xk[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xk[d] = ev.x[(size_t)d * npad + ic];
    const double tk = ev.t[ic];
    const double wk = w[ic];

    const int wave_first = __builtin_amdgcn_readfirstlane(rb * kRowsPerBlock + (tid & ~63));
    const bool wave_live = wave_first < n;
    const double t_lo = ev.t[min(wave_first, n - 1)];
    const double t_hi = ev.t[min(wave_first + 63, n - 1)];

    double acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] = 0.0;
    const int j_begin = c * g.chunk_len;
    const int j_end = min(n, j_begin + g.chunk_len);
    for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
        const int cnt = min(kTile, j_end - j0);
        __syncthreads();
        if (tid < cnt) {
#pragma unroll
            for (int d = 0; d < D; ++d) tile[tid * LS + d] = ev.x[(size_t)d * npad + j0 + tid];
            tile[tid * LS + D] = ev.t[j0 + tid];
            tile[tid * LS + D + 1] = w[j0 + tid];
        }
        __syncthreads();
        if (wave_live) {
            const double tf = tile[D], tl = tile[(cnt - 1) * LS + D];
            int smode = tl < t_lo ? kLBefore : (tf > t_hi ? kLAfter : kLMixed);
            int dob = 1;
            if (g.skip_zero) {
                // smallest |dt| between the wavefront's rows and the tile; skip only exact zeros
                const double gap = tl < t_lo ? t_lo - tl : (tf > t_hi ? tf - t_hi : 0.0);
                if (-cBt * gap * gap >= kZeroExpArg) dob = 0;
                if (smode != kLMixed && omega >= 0.0 && omega * gap >= kZeroExpArg) smode = kLNone;
            }
            const int mode = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
            switch (mode) {
            case 4 + kLNone: locgrad_tile<D, true, kLNone>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case 4 + kLBefore: locgrad_tile<D, true, kLBefore>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case 4 + kLAfter: locgrad_tile<D, true, kLAfter>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case 4 + kLMixed: locgrad_tile<D, true, kLMixed>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case kLBefore: locgrad_tile<D, false, kLBefore>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case kLAfter: locgrad_tile<D, false, kLAfter>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            case kLMixed: locgrad_tile<D, false, kLMixed>(tile, cnt, xk, tk, wk, cBx, cBt, cSx, omega, fB, fS, acc); break;
            default: break; // every term of the tile is exactly zero
            }
        }
    }
    double *out = part + (size_t)c * D * npad;
#pragma unroll
    for (int d = 0; d < D; ++d) out[(size_t)d * npad + i] = acc[d];
}

// Fixed-order sum of the chunk partials; out[perm[i] * D + d] in the caller's event order.
__global__ __launch_bounds__(kRowsPerBlock) void locgrad_finalize_kernel(Events ev, PairGrid g,
                                                                         const double *__restrict__ part,
                                                                         double *__restrict__ out) {
    const int i = blockIdx.x * kRowsPerBlock + threadIdx.x;
    const int n = ev.n, npad = ev.npad, D = ev.dim;
    if (i >= n) return;
    const int u = ev.perm[i];
    for (int d = 0; d < D; ++d) {
        double s = 0.0;
        #pragma unroll 4
        for (int c = 0; c < g.chunks; ++c) s += part[((size_t)c * D + d) * npad + i];
        out[(size_t)u * D + d] = s;
    }
}

// ---- symmetric form (D <= 3): each background exp serves both events of a pair ----
//
// Same work decomposition, queue and partial layout as the symmetric likelihood pass with
// gradient (plan_sym(..., grad = true): 6 row and 3 column accumulators, of which the first D
// are used). A row block is paired with earlier columns only; for a row i and an earlier column
// j, f = fB eB_ij (w_i + w_j) + fS eS_ij w_i (eS_ji = 0), and d ll / d x_i gets -f (x_i - x_j),
// d ll / d x_j gets +f (x_i - x_j) (column accumulators that move one lane per step, ds_bpermute;
// this pass is bound by its natural-form exp, and the likelihood pass's LDS column slots measured
// no faster here).
// The diagonal block is paired in both orders, each order adding to its row only.
// The chunk subtiles' exponentials from the LDS table (exp2_neg_tab): each exponent scaled by
// 256 log2(e), formed from double-double constants (the scale of tauX^2 / 2, tauT^2 / 2,
// sigma^2 / 2 and omega the finalize's) on the natural differences. The excitation terms take it
// for omega >= 0 only (exc; their exponent has one sign then), else exp_fp64.
struct LocTab {
    const double *tab;
    double2 b0, b1, s0, s1; // 256 log2(e) x (tauX^2 / 2, tauT^2 / 2, sigma^2 / 2, omega)
    bool exc;
};
__device__ __forceinline__ double2 scale_256log2e(double v) {
    const double kh = 256.0 * kLog2eHi, kl = 256.0 * kLog2eLo;
    const double hi = v * kh;
    return make_double2(hi, __builtin_fma(v, kh, -hi) + v * kl);
}
__device__ __forceinline__ double tab_exp(double2 c0, double x0, double2 c1, double x1, const double *tab) {
    return exp2_neg_tab(__builtin_fma(c0.x, x0, __builtin_fma(c0.y, x0, __builtin_fma(c1.x, x1, c1.y * x1))), tab);
}

template <int D, int A, bool DOB, int SMODE, bool TAIL>
__device__ __forceinline__ void loc_subtile(const double *__restrict__ sub, int lane, const double (&xi)[A][D],
                                            const double (&ti)[A], const double (&wi)[A], const double (&vm)[A],
                                            double cSx, double nomega, double fB, double fS, const LocTab &lt,
                                            double (&acc)[A][D], double (&ca)[D]) {
    const int src = ((lane + 1) & 63) << 2;
#pragma unroll 2
    for (int s = 0; s < 64; ++s) {
        // structure-of-arrays subtile: x (D arrays), t, w; component d of entry u at sub[d * 128 + u]
        double q[D + 2];
#pragma unroll
        for (int d = 0; d < D + 2; ++d) q[d] = sub[d * 128 + lane + s];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            double dx[D];
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                dx[d] = xi[a][d] - q[d];
                r2 = __builtin_fma(dx[d], dx[d], r2);
            }
            const double dt = ti[a] - q[D];
            double f = 0.0;
            if (DOB) f = fB * tab_exp(lt.b0, r2, lt.b1, dt * dt, lt.tab) * (wi[a] + q[D + 1]);
            if (SMODE != kSNone) {
                // (a pair with dt <= 0 may give +inf or a wrong value: masked below)
                double es = lt.exc ? tab_exp(lt.s0, r2, lt.s1, dt, lt.tab) : exp_fp64(__builtin_fma(cSx, r2, nomega * dt));
                if (SMODE == kSMixed) es = dt > 0.0 ? es : 0.0;
                f = __builtin_fma(fS * es, wi[a], f);
            }
            const double fc = TAIL ? f * vm[a] : f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                acc[a][d] = __builtin_fma(-f, dx[d], acc[a][d]);
                ca[d] = __builtin_fma(fc, dx[d], ca[d]);
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) ca[d] = lane_shift_down(ca[d], src);
    }
}

template <int D, int A, bool TAIL>
__device__ __forceinline__ void loc_dispatch(int mode, const double *__restrict__ sub, int lane,
                                             const double (&xi)[A][D], const double (&ti)[A], const double (&wi)[A],
                                             const double (&vm)[A], double cSx, double nomega, double fB,
                                             double fS, const LocTab &lt, double (&acc)[A][D], double (&ca)[D]) {
    switch (mode) {
    case 4 + kSNone:
        loc_subtile<D, A, true, kSNone, TAIL>(sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc, ca);
        break;
    case 4 + kSFull:
        loc_subtile<D, A, true, kSFull, TAIL>(sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc, ca);
        break;
    case 4 + kSMixed:
        loc_subtile<D, A, true, kSMixed, TAIL>(sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc, ca);
        break;
    case kSFull:
        loc_subtile<D, A, false, kSFull, TAIL>(sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc, ca);
        break;
    case kSMixed:
        loc_subtile<D, A, false, kSMixed, TAIL>(sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc, ca);
        break;
    default: break; // every term of the subtile is exactly zero
    }
}

template <int D, int A>
__global__ __launch_bounds__(kSymWaves * 64) void sym_locgrad_kernel(SymLaunch L, const double *__restrict__ params,
                                                                     const double *__restrict__ w,
                                                                     double *__restrict__ rowpart,
                                                                     double *__restrict__ colpart) {
    constexpr int LS = LocStride<D>::value;
    constexpr int NR = 6, NC = 3; // partial layout of the gradient plan
    constexpr int R = kSymWaves * 64 * A;
    // diagonal items: kTile events x LS doubles; chunk items: 4 subtiles x (D + 2) x 128 doubles
    __shared__ double tile[(2 * kTile * (D + 2) > kTile * LS) ? 2 * kTile * (D + 2) : kTile * LS];
    __shared__ double colbuf[kSymWaves * D * kTile];
    __shared__ double etab[kExp2TabSize];
    __shared__ int s_item;

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the table is complete before the first item: the queue loop's barrier follows
    fill_exp2_table(etab, tid, kSymWaves * 64);
    for (;;) {
        const int idx = queue_next(L, &s_item);
        if (idx >= L.items) break;

        const int4 it = L.item[idx];
        const SymData &ds = L.data[it.x];
        const int P = it.y, chunk = it.z, k = it.w;
        const double *pk = params + 6 * (size_t)(ds.param0 + k);
        const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3], theta = pk[4], mu0 = pk[5];
        const double cBx = -0.5 * tauX * tauX, cBt = -0.5 * tauT * tauT, cSx = -0.5 * sigma * sigma;
        const double nomega = -omega;
        LocTab lt;
        lt.tab = etab;
        lt.b0 = scale_256log2e(-cBx);
        lt.b1 = scale_256log2e(-cBt);
        lt.s0 = scale_256log2e(-cSx);
        lt.s1 = scale_256log2e(omega);
        lt.exc = omega >= 0.0;
        const double fB = mu0 * ipow(tauX, D) * tauT * 0.15915494309189533577 * tauX * tauX;
        const double fS = theta * omega * ipow(sigma, D) * 0.39894228040143267794 * sigma * sigma;
        const double *__restrict__ ex = ds.x;
        const double *__restrict__ et = ds.t;
        const int n = ds.n, npad = ds.npad;
        const int P0 = P * R;

        const int wrow0 = __builtin_amdgcn_readfirstlane(P0 + wv * 64 * A);
        double xi[A][D], ti[A], wi[A], vm[A];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const int r = wrow0 + a * 64 + lane;
            const int rc = min(r, n - 1);
#pragma unroll
            for (int d = 0; d < D; ++d) xi[a][d] = ex[(size_t)d * npad + rc];
            ti[a] = et[rc];
            wi[a] = w[rc];
            vm[a] = r < n ? 1.0 : 0.0;
        }
        const bool wave_live = wrow0 < n;
        const double t_lo = et[min(wrow0, n - 1)];

        double acc[A][D];
#pragma unroll
        for (int a = 0; a < A; ++a)
#pragma unroll
            for (int d = 0; d < D; ++d) acc[a][d] = 0.0;

        int slot;
        if (chunk < 0) {
            slot = ds.max_chunks;
            const int c1 = min(P0 + R, n);
            for (int j0 = P0; j0 < c1; j0 += kTile) {
                const int cnt = min(kTile, c1 - j0);
                __syncthreads();
                if (tid < cnt) {
#pragma unroll
                    for (int d = 0; d < D; ++d) tile[tid * LS + d] = ex[(size_t)d * npad + j0 + tid];
                    tile[tid * LS + D] = et[j0 + tid];
                    tile[tid * LS + D + 1] = w[j0 + tid];
                }
                __syncthreads();
                if (wave_live) {
#pragma unroll
                    for (int a = 0; a < A; ++a)
                        locgrad_tile<D, true, kLMixed>(tile, cnt, xi[a], ti[a], wi[a], cBx, cBt, cSx, omega, fB, fS,
                                                      acc[a]);
                }
            }
        } else {
            slot = chunk;
            const int c0 = chunk * ds.L;
            const int c1 = min(c0 + ds.L, P0);
            const bool tail = P0 + R > n;
            double *cp = colpart + ds.col0 + (size_t)k * ds.col_kstride + (size_t)NC * R * ((size_t)P * (P - 1) / 2);
            const double t_blk = et[min(P0, n - 1)];
            for (int j0 = c0; j0 < c1; j0 += kTile) {
                // c1 - c0 and P0 are multiples of kTile: every tile here is full.
                if (L.skip_zero) {
                    // workgroup-wide skip when no term survives for the block's first row
                    const double tl = et[j0 + kTile - 1];
                    const double gap = t_blk - tl;
                    if (tl < t_blk && -cBt * gap * gap >= kZeroExpArg && omega >= 0.0 && omega * gap >= kZeroExpArg) {
#pragma unroll
                        for (int c = 0; c < D; ++c) cp[(size_t)c * P0 + j0 + tid] = 0.0;
                        continue;
                    }
                }
                __syncthreads();
                {
                    const int e0 = (tid >> 6) * (D + 2) * 128 + (tid & 63);
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const double v = ex[(size_t)d * npad + j0 + tid];
                        tile[e0 + d * 128] = v;
                        tile[e0 + d * 128 + 64] = v;
                    }
                    const double tv = et[j0 + tid], wv2 = w[j0 + tid];
                    tile[e0 + D * 128] = tv;
                    tile[e0 + D * 128 + 64] = tv;
                    tile[e0 + (D + 1) * 128] = wv2;
                    tile[e0 + (D + 1) * 128 + 64] = wv2;
                }
                __syncthreads();
                if (wave_live) {
                    const double tl = tile[(3 * (D + 2) + D) * 128 + 63];
                    int smode = tl < t_lo ? kSFull : kSMixed;
                    int dob = 1;
                    if (L.skip_zero) {
                        const double gap = tl < t_lo ? t_lo - tl : 0.0;
                        if (-cBt * gap * gap >= kZeroExpArg) dob = 0;
                        if (smode == kSFull && omega >= 0.0 && omega * gap >= kZeroExpArg) smode = kSNone;
                    }
                    const int mode = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
                    for (int sb = 0; sb < 4; ++sb) {
                        double ca[D];
#pragma unroll
                        for (int c = 0; c < D; ++c) ca[c] = 0.0;
                        const double *sub = tile + sb * (D + 2) * 128;
                        if (tail)
                            loc_dispatch<D, A, true>(mode, sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc,
                                                     ca);
                        else
                            loc_dispatch<D, A, false>(mode, sub, lane, xi, ti, wi, vm, cSx, nomega, fB, fS, lt, acc,
                                                      ca);
#pragma unroll
                        for (int c = 0; c < D; ++c) colbuf[(wv * D + c) * kTile + sb * 64 + lane] = ca[c];
                    }
                } else {
#pragma unroll
                    for (int sb = 0; sb < 4; ++sb)
#pragma unroll
                        for (int c = 0; c < D; ++c) colbuf[(wv * D + c) * kTile + sb * 64 + lane] = 0.0;
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    double v = colbuf[(0 * D + c) * kTile + tid];
#pragma unroll
                    for (int ww = 1; ww < kSymWaves; ++ww) v += colbuf[(ww * D + c) * kTile + tid];
                    cp[(size_t)c * P0 + j0 + tid] = v;
                }
            }
        }
        const size_t rpad = (size_t)ds.nb * R;
        double *rp = rowpart + ds.row0 + (size_t)k * ds.row_kstride + (size_t)slot * NR * rpad;
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const int r = wrow0 + a * 64 + lane;
#pragma unroll
            for (int d = 0; d < D; ++d) rp[(size_t)d * rpad + r] = acc[a][d];
        }
    }
}

// Row sums of the symmetric location-gradient pass (row chunks, diagonal block, column side of
// every later block, in a fixed order) scattered to out[perm[i] * D + d] (dataset 0, vector 0).
template <int D>
__global__ __launch_bounds__(kRowsPerBlock) void sym_locgrad_finalize_kernel(SymLaunch L,
                                                                             const double *__restrict__ rowpart,
                                                                             const double *__restrict__ colpart,
                                                                             double *__restrict__ out) {
    constexpr int NR = 6, NC = 3;
    const int4 it = L.fin_item[blockIdx.x];
    if (it.x != 0 || it.y != 0) return;
    const SymData &ds = L.data[0];
    const int i = it.z * kRowsPerBlock + threadIdx.x;
    const int n = ds.n, R = L.R;
    if (i >= n) return;
    const int P = i / R;
    const int nch = (P * R + ds.L - 1) / ds.L;
    const size_t rpad = (size_t)ds.nb * R;
    const double *rbase = rowpart + ds.row0;
    double a[D];
#pragma unroll
    for (int d = 0; d < D; ++d) a[d] = 0.0;
    #pragma unroll 4
    for (int c = 0; c < nch; ++c) {
        const double *rp = rbase + (size_t)c * NR * rpad;
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] += rp[(size_t)d * rpad + i];
    }
    {
        const double *rp = rbase + (size_t)ds.max_chunks * NR * rpad;
#pragma unroll
        for (int d = 0; d < D; ++d) a[d] += rp[(size_t)d * rpad + i];
    }
    double cs[D];
#pragma unroll
    for (int d = 0; d < D; ++d) cs[d] = 0.0;
    const double *cbase = colpart + ds.col0;
    #pragma unroll 4
    for (int Q = P + 1; Q < ds.nb; ++Q) {
        const double *cp = cbase + (size_t)NC * R * ((size_t)Q * (Q - 1) / 2);
        const size_t Q0 = (size_t)Q * R;
#pragma unroll
        for (int d = 0; d < D; ++d) cs[d] += cp[(size_t)d * Q0 + i];
    }
    const int u = ds.perm[i];
#pragma unroll
    for (int d = 0; d < D; ++d) out[(size_t)u * D + d] = a[d] + cs[d];
}

template <int D>
static void launch_sym_locgrad_d(const SymLaunch &L, const double *params, const double *w, double *rowpart,
                                 double *colpart, double *out, hipStream_t s) {
    if (L.A == 2)
        hipLaunchKernelGGL((sym_locgrad_kernel<D, 2>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s, L, params,
                           w, rowpart, colpart);
    else
        hipLaunchKernelGGL((sym_locgrad_kernel<D, 1>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s, L, params,
                           w, rowpart, colpart);
    hipLaunchKernelGGL((sym_locgrad_finalize_kernel<D>), dim3((unsigned)L.fin_items), dim3(kRowsPerBlock), 0, s, L,
                       rowpart, colpart, out);
}

void launch_sym_locgrad(const SymLaunch &L, const double *params, const double *w, double *rowpart,
                        double *colpart, double *out, hipStream_t s) {
    if (L.A != 1 && L.A != 2) throw std::invalid_argument("symmetric pass is built for 1 or 2 rows per lane");
    if (L.R != kSymWaves * 64 * L.A) throw std::invalid_argument("symmetric pass: block rows do not match");
    if (L.grid < 1 || L.items < 1) throw std::invalid_argument("symmetric pass: empty launch");
    switch (L.dim) {
    case 1: launch_sym_locgrad_d<1>(L, params, w, rowpart, colpart, out, s); break;
    case 2: launch_sym_locgrad_d<2>(L, params, w, rowpart, colpart, out, s); break;
    case 3: launch_sym_locgrad_d<3>(L, params, w, rowpart, colpart, out, s); break;
    default: throw std::invalid_argument("symmetric pass is built for dimensions 1..3");
    }
}

template <int D>
static void launch_locgrad_d(const Events &ev, const double *params, const double *w, const PairGrid &g,
                             double *part, hipStream_t s) {
    hipLaunchKernelGGL((locgrad_kernel<D>), dim3((unsigned)(g.row_blocks * g.chunks)), dim3(kRowsPerBlock), 0, s,
                       ev, params, w, g, part);
}

void launch_inv_rate(const Events &ev, const double *params, const PairGrid &g, const double *rate_part, double *w,
                     hipStream_t s) {
    hipLaunchKernelGGL(inv_rate_kernel, dim3((unsigned)(ev.npad / kRowsPerBlock)), dim3(kRowsPerBlock), 0, s, ev,
                       params, g, rate_part, w);
}

void launch_locgrad(const Events &ev, const double *params, const double *w, const PairGrid &g, double *part,
                    double *out, hipStream_t s) {
    switch (ev.dim) {
    case 1: launch_locgrad_d<1>(ev, params, w, g, part, s); break;
    case 2: launch_locgrad_d<2>(ev, params, w, g, part, s); break;
    case 3: launch_locgrad_d<3>(ev, params, w, g, part, s); break;
    case 4: launch_locgrad_d<4>(ev, params, w, g, part, s); break;
    case 5: launch_locgrad_d<5>(ev, params, w, g, part, s); break;
    case 6: launch_locgrad_d<6>(ev, params, w, g, part, s); break;
    case 7: launch_locgrad_d<7>(ev, params, w, g, part, s); break;
    case 8: launch_locgrad_d<8>(ev, params, w, g, part, s); break;
    default: throw std::invalid_argument("embedding dimension must be 1..8");
    }
    hipLaunchKernelGGL(locgrad_finalize_kernel, dim3((unsigned)(ev.npad / kRowsPerBlock)), dim3(kRowsPerBlock), 0, s,
                       ev, g, part, out);
}

} // namespace kern
} // namespace hph
