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

// Stride (doubles) of one column in the LDS tile: D coordinates, t, w, padded to 16 bytes.
template <int D> struct LocStride {
    static constexpr int value = (D + 3) & ~1;
};

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
        for (int c = 0; c < g.chunks; ++c) s += part[((size_t)c * D + d) * npad + i];
        out[(size_t)u * D + d] = s;
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
