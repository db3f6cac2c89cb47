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
// Column sums inside a wavefront use a rotation: over 64 steps lane l pairs its rows with column
// (l + s) mod 64 of a 64-column subtile and adds into a 3-double accumulator that moves one lane
// down after every step (ds_bpermute), so after 64 steps lane l holds the sums of column l.
// The subtile is stored twice in LDS, as structure of arrays, so the per-lane address (l + s)
// needs no wrap and consecutive lanes read consecutive doubles (no bank conflicts).
#include "hph_device.h"

#include <stdexcept>
#include <vector>

namespace hph {
namespace kern {

constexpr int kSymWaves = 4;

__device__ __forceinline__ double lane_shift_down(double v, int byte_addr) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)u);
    const int hi = __builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// One 64-column subtile against this lane's A rows, rotating the column accumulators ca.
// DOB: form the background term; SMODE: excitation absent / unmasked / masked; TAIL: the block
// has rows past the last event, whose column contributions are masked by vm; F32: FP32-bulk
// background exponential (cBx, cBt then carry the log2 e scale).
template <int D, bool GRAD, int A, bool DOB, int SMODE, bool TAIL, bool F32>
__device__ __forceinline__ void sym_subtile(const double *__restrict__ sub, int lane, const double (&xi)[A][D],
                                            const double (&ti)[A], const double (&vm)[A], double cBx, double cBt,
                                            double cSx, double nomega, double (&acc)[A][6],
                                            double (&ca)[GRAD ? 3 : 1]) {
    constexpr int NC = GRAD ? 3 : 1;
    const int src = ((lane + 1) & 63) << 2;
#pragma unroll 2
    for (int s = 0; s < 64; ++s) {
        // structure-of-arrays subtile: component d of entry u at sub[d * 128 + u]
        double q[D + 1];
#pragma unroll
        for (int d = 0; d <= D; ++d) q[d] = sub[d * 128 + lane + s];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            double r2;
            {
                const double d0 = xi[a][0] - q[0];
                r2 = d0 * d0;
            }
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const double dd = xi[a][d] - q[d];
                r2 = __builtin_fma(dd, dd, r2);
            }
            const double dt = ti[a] - q[D];
            if (DOB) {
                const double dt2 = dt * dt;
                const double eb = background_exp<F32>(cBx, cBt, r2, dt2);
                acc[a][0] += eb;
                if (GRAD) {
                    acc[a][1] = __builtin_fma(eb, r2, acc[a][1]);
                    acc[a][2] = __builtin_fma(eb, dt2, acc[a][2]);
                }
                const double ec = TAIL ? eb * vm[a] : eb;
                ca[0] += ec;
                if (GRAD) {
                    ca[1] = __builtin_fma(ec, r2, ca[1]);
                    ca[2] = __builtin_fma(ec, dt2, ca[2]);
                }
            }
            if (SMODE != kSNone) {
                double es = exp_fp64(__builtin_fma(cSx, r2, nomega * dt));
                if (SMODE == kSMixed) es = dt > 0.0 ? es : 0.0;
                acc[a][3] += es;
                if (GRAD) {
                    acc[a][4] = __builtin_fma(es, r2, acc[a][4]);
                    acc[a][5] = __builtin_fma(es, dt, acc[a][5]);
                }
            }
        }
        if (DOB) {
#pragma unroll
            for (int c = 0; c < NC; ++c) ca[c] = lane_shift_down(ca[c], src);
        }
    }
}

// One diagonal-block tile (plain layout, broadcast reads) against one row: ordered pairs with
// the dt > 0 mask.
template <int D, bool GRAD, bool F32>
__device__ __forceinline__ void diag_pass(const double *__restrict__ tile, int cnt, const double (&xi)[D], double ti,
                                          double cBx, double cBt, double cSx, double nomega, double (&acc)[6]) {
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
        const double dt2 = dt * dt;
        const double eb = background_exp<F32>(cBx, cBt, r2, dt2);
        acc[0] += eb;
        if (GRAD) {
            acc[1] = __builtin_fma(eb, r2, acc[1]);
            acc[2] = __builtin_fma(eb, dt2, acc[2]);
        }
        double es = exp_fp64(__builtin_fma(cSx, r2, nomega * dt));
        es = dt > 0.0 ? es : 0.0;
        acc[3] += es;
        if (GRAD) {
            acc[4] = __builtin_fma(es, r2, acc[4]);
            acc[5] = __builtin_fma(es, dt, acc[5]);
        }
    }
}

template <int D, bool GRAD, int A, bool TAIL, bool F32>
__device__ __forceinline__ void sym_dispatch(int mode, const double *__restrict__ sub, int lane,
                                             const double (&xi)[A][D], const double (&ti)[A], const double (&vm)[A],
                                             double cBx, double cBt, double cSx, double nomega, double (&acc)[A][6],
                                             double (&ca)[GRAD ? 3 : 1]) {
    switch (mode) {
    case 4 + kSNone:
        sym_subtile<D, GRAD, A, true, kSNone, TAIL, F32>(sub, lane, xi, ti, vm, cBx, cBt, cSx, nomega, acc, ca);
        break;
    case 4 + kSFull:
        sym_subtile<D, GRAD, A, true, kSFull, TAIL, F32>(sub, lane, xi, ti, vm, cBx, cBt, cSx, nomega, acc, ca);
        break;
    case 4 + kSMixed:
        sym_subtile<D, GRAD, A, true, kSMixed, TAIL, F32>(sub, lane, xi, ti, vm, cBx, cBt, cSx, nomega, acc, ca);
        break;
    case kSFull:
        sym_subtile<D, GRAD, A, false, kSFull, TAIL, F32>(sub, lane, xi, ti, vm, cBx, cBt, cSx, nomega, acc, ca);
        break;
    case kSMixed:
        sym_subtile<D, GRAD, A, false, kSMixed, TAIL, F32>(sub, lane, xi, ti, vm, cBx, cBt, cSx, nomega, acc, ca);
        break;
    default: break; // every term of the subtile is exactly zero
    }
}

// Grid: x = work item (row block, chunk of earlier columns | diagonal), y = parameter vector k.
template <int D, bool GRAD, int A, bool F32>
__global__ __launch_bounds__(kSymWaves * 64) void sym_kernel(Events ev, const double *__restrict__ params,
                                                             SymGrid g, double *__restrict__ rowpart,
                                                             double *__restrict__ colpart) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    // diagonal items: kTile events x ST doubles; chunk items: 4 subtiles x (D + 1) x 128 doubles
    __shared__ double tile[(2 * kTile * (D + 1) > kTile * ST) ? 2 * kTile * (D + 1) : kTile * ST];
    __shared__ double colbuf[kSymWaves * NC * kTile];

    const int2 it = g.item[blockIdx.x];
    const int P = it.x, chunk = it.y;
    const int k = blockIdx.y;
    const double *pk = params + 6 * k;
    const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3];
    const double cBx = -0.5 * tauX * tauX, cBt = -0.5 * tauT * tauT;
    const double cSx = -0.5 * sigma * sigma, nomega = -omega;
    // background-exponential constants (log2 e scaled for the FP32-bulk v_exp_f32)
    const double eBx = F32 ? cBx * kLog2e : cBx, eBt = F32 ? cBt * kLog2e : cBt;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = ev.n, npad = ev.npad;
    const int R = g.R;
    const int P0 = P * R;
    const int wrow0 = __builtin_amdgcn_readfirstlane(P0 + w * 64 * A);
    double xi[A][D], ti[A], vm[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int r = wrow0 + a * 64 + lane;
        const int rc = min(r, n - 1);
#pragma unroll
        for (int d = 0; d < D; ++d) xi[a][d] = ev.x[(size_t)d * npad + rc];
        ti[a] = ev.t[rc];
        vm[a] = r < n ? 1.0 : 0.0;
    }
    const bool wave_live = wrow0 < n;
    const double t_lo = ev.t[min(wrow0, n - 1)];
    const double t_hi = ev.t[min(wrow0 + 64 * A - 1, n - 1)];
    (void)t_hi;

    double acc[A][6];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[a][q] = 0.0;

    int slot;
    if (chunk < 0) {
        slot = g.max_chunks;
        const int c1 = min(P0 + R, n);
        for (int j0 = P0; j0 < c1; j0 += kTile) {
            const int cnt = min(kTile, c1 - j0);
            __syncthreads();
            if (tid < cnt) {
#pragma unroll
                for (int d = 0; d < D; ++d) tile[tid * ST + d] = ev.x[(size_t)d * npad + j0 + tid];
                tile[tid * ST + D] = ev.t[j0 + tid];
            }
            __syncthreads();
            if (wave_live) {
#pragma unroll
                for (int a = 0; a < A; ++a) diag_pass<D, GRAD, F32>(tile, cnt, xi[a], ti[a], eBx, eBt, cSx, nomega, acc[a]);
            }
        }
    } else {
        slot = chunk;
        const int c0 = chunk * g.L;
        const int c1 = min(c0 + g.L, P0);
        const bool tail = P0 + R > n;
        double *cp = colpart + (size_t)k * g.col_stride + (size_t)NC * R * ((size_t)P * (P - 1) / 2);
        for (int j0 = c0; j0 < c1; j0 += kTile) {
            // c1 - c0 and P0 are multiples of kTile: every tile here is full.
            __syncthreads();
            {
                // subtile sb = tid / 64 holds columns j0 + 64 sb .. + 63 as D + 1 arrays of 128
                // doubles (each 64-column run stored twice); lanes then read consecutive doubles.
                const int e0 = (tid >> 6) * (D + 1) * 128 + (tid & 63);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double v = ev.x[(size_t)d * npad + j0 + tid];
                    tile[e0 + d * 128] = v;
                    tile[e0 + d * 128 + 64] = v;
                }
                const double tv = ev.t[j0 + tid];
                tile[e0 + D * 128] = tv;
                tile[e0 + D * 128 + 64] = tv;
            }
            __syncthreads();
            if (wave_live) {
                // columns are earlier than the rows: tl <= t_lo; equal times need the mask
                const double tl = tile[(3 * (D + 1) + D) * 128 + 63];
                int smode = tl < t_lo ? kSFull : kSMixed;
                int dob = 1;
                if (g.skip_zero) {
                    const double gap = tl < t_lo ? t_lo - tl : 0.0;
                    if (-cBt * gap * gap >= g.zero_arg_b) dob = 0;
                    if (smode == kSFull && nomega <= 0.0 && -nomega * gap >= g.zero_arg_s) smode = kSNone;
                }
                const int mode = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
                for (int sb = 0; sb < 4; ++sb) {
                    double ca[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) ca[c] = 0.0;
                    const double *sub = tile + sb * (D + 1) * 128;
                    if (tail)
                        sym_dispatch<D, GRAD, A, true, F32>(mode, sub, lane, xi, ti, vm, eBx, eBt, cSx, nomega, acc, ca);
                    else
                        sym_dispatch<D, GRAD, A, false, F32>(mode, sub, lane, xi, ti, vm, eBx, eBt, cSx, nomega, acc, ca);
#pragma unroll
                    for (int c = 0; c < NC; ++c) colbuf[(w * NC + c) * kTile + sb * 64 + lane] = ca[c];
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
        }
    }
    const size_t rpad = (size_t)g.nb * R;
    double *rp = rowpart + ((size_t)k * (g.max_chunks + 1) + slot) * NR * rpad;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int r = wrow0 + a * 64 + lane;
        if (GRAD) {
#pragma unroll
            for (int q = 0; q < 6; ++q) rp[(size_t)q * rpad + r] = acc[a][q];
        } else {
            rp[r] = acc[a][0];
            rp[rpad + r] = acc[a][3];
        }
    }
}

// Row sums = row-side chunks + diagonal block + column side of every later block.
template <bool GRAD>
__global__ __launch_bounds__(kRowsPerBlock) void sym_finalize_kernel(Events ev, const double *__restrict__ params,
                                                                     SymGrid g, const double *__restrict__ rowpart,
                                                                     const double *__restrict__ colpart,
                                                                     double *__restrict__ block_out,
                                                                     double *__restrict__ probs) {
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    __shared__ double2 sh[kRowsPerBlock];
    const int k = blockIdx.y;
    const int fb = blockIdx.x;
    const int i = fb * kRowsPerBlock + threadIdx.x;
    const int n = ev.n;
    const int R = g.R;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const int P = i / R;
        const int nch = (P * R + g.L - 1) / g.L;
        const size_t rpad = (size_t)g.nb * R;
        double a6[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) a6[q] = 0.0;
        for (int c = 0; c < nch; ++c) {
            const double *rp = rowpart + ((size_t)k * (g.max_chunks + 1) + c) * NR * rpad;
#pragma unroll
            for (int q = 0; q < NR; ++q) a6[q] += rp[(size_t)q * rpad + i];
        }
        {
            const double *rp = rowpart + ((size_t)k * (g.max_chunks + 1) + g.max_chunks) * NR * rpad;
#pragma unroll
            for (int q = 0; q < NR; ++q) a6[q] += rp[(size_t)q * rpad + i];
        }
        double cs[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) cs[c] = 0.0;
        for (int Q = P + 1; Q < g.nb; ++Q) {
            const double *cp = colpart + (size_t)k * g.col_stride + (size_t)NC * R * ((size_t)Q * (Q - 1) / 2);
            const size_t Q0 = (size_t)Q * R;
#pragma unroll
            for (int c = 0; c < NC; ++c) cs[c] += cp[(size_t)c * Q0 + i];
        }
        double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (GRAD) {
            s[0] = a6[0] + cs[0];
            s[1] = a6[1] + cs[NC > 1 ? 1 : 0];
            s[2] = a6[2] + cs[NC > 2 ? 2 : 0];
            s[3] = a6[3];
            s[4] = a6[4];
            s[5] = a6[5];
        } else {
            s[0] = a6[0] + cs[0];
            s[3] = a6[NR - 1];
        }
        const double p = row_terms<GRAD>(s, params + 6 * k, ev.dim, ev.t[i], ev.t_last, v);
        if (probs != nullptr && k == 0) probs[ev.perm[i]] = p;
    }
    store_block_sums<GRAD>(v, sh, block_out, k, (n + kRowsPerBlock - 1) / kRowsPerBlock, fb);
}

// ---------------- host side ----------------

static long long sym_items_for(int nb, int R, int L) {
    long long items = nb; // diagonal items
    for (int P = 1; P < nb; ++P) items += ((long long)P * R + L - 1) / L;
    return items;
}

SymPlan plan_sym_grid(int n, int dim, int K, int target_blocks, int rows_per_lane) {
    if (rows_per_lane != 1 && rows_per_lane != 2) throw std::invalid_argument("rows per lane must be 1 or 2");
    SymPlan plan;
    SymGrid &g = plan.grid;
    g.A = rows_per_lane;
    g.R = kSymWaves * 64 * g.A;
    g.nb = (n + g.R - 1) / g.R;
    const int max_len = ((g.nb > 1 ? (g.nb - 1) * g.R : g.R) + kTile - 1) / kTile * kTile;
    int L = max_len;
    while (L > kTile && (long long)K * sym_items_for(g.nb, g.R, L) < target_blocks) L = ((L / 2 + kTile - 1) / kTile) * kTile;
    g.L = L;
    g.max_chunks = g.nb > 1 ? ((g.nb - 1) * g.R + L - 1) / L : 1;
    const int NC = 3;
    g.col_stride = (long long)NC * g.R * ((long long)g.nb * (g.nb - 1) / 2);
    g.skip_zero = 1;
    g.zero_arg_b = kZeroExpArg;
    g.zero_arg_s = kZeroExpArg;
    plan.items.clear();
    for (int P = g.nb - 1; P >= 1; --P) {
        const int nch = (P * g.R + L - 1) / L;
        for (int c = 0; c < nch; ++c) plan.items.push_back(make_int2(P, c));
    }
    for (int P = g.nb - 1; P >= 0; --P) plan.items.push_back(make_int2(P, -1));
    g.items = (int)plan.items.size();
    g.item = nullptr;
    (void)dim;
    return plan;
}

size_t sym_rowpart_doubles(const SymGrid &g, int K, bool grad) {
    return (size_t)K * (g.max_chunks + 1) * (grad ? 6 : 2) * (size_t)g.nb * g.R;
}

size_t sym_colpart_doubles(const SymGrid &g, int K, bool grad) {
    (void)grad;
    return (size_t)K * (size_t)g.col_stride;
}

template <int D, bool GRAD, bool F32>
static void launch_sym_d(const Events &ev, const double *params, int K, const SymGrid &g, double *rowpart,
                         double *colpart, hipStream_t s) {
    dim3 grid((unsigned)g.items, (unsigned)K);
    if (g.A == 2)
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 2, F32>), grid, dim3(kSymWaves * 64), 0, s, ev, params, g, rowpart,
                           colpart);
    else
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 1, F32>), grid, dim3(kSymWaves * 64), 0, s, ev, params, g, rowpart,
                           colpart);
}

template <bool F32>
static void launch_sym_p(const Events &ev, const double *params, int K, const SymGrid &g, bool grad,
                         double *rowpart, double *colpart, hipStream_t s) {
    switch (ev.dim * 2 + (grad ? 1 : 0)) {
    case 2: launch_sym_d<1, false, F32>(ev, params, K, g, rowpart, colpart, s); break;
    case 3: launch_sym_d<1, true, F32>(ev, params, K, g, rowpart, colpart, s); break;
    case 4: launch_sym_d<2, false, F32>(ev, params, K, g, rowpart, colpart, s); break;
    case 5: launch_sym_d<2, true, F32>(ev, params, K, g, rowpart, colpart, s); break;
    case 6: launch_sym_d<3, false, F32>(ev, params, K, g, rowpart, colpart, s); break;
    case 7: launch_sym_d<3, true, F32>(ev, params, K, g, rowpart, colpart, s); break;
    default: throw std::invalid_argument("symmetric pass is built for dimensions 1..3");
    }
}

void launch_sym(const Events &ev, const double *params, int K, const SymGrid &g, bool grad, bool fp32,
                double *rowpart, double *colpart, hipStream_t s) {
    if (g.A != 1 && g.A != 2) throw std::invalid_argument("symmetric pass is built for 1 or 2 rows per lane");
    if (fp32)
        launch_sym_p<true>(ev, params, K, g, grad, rowpart, colpart, s);
    else
        launch_sym_p<false>(ev, params, K, g, grad, rowpart, colpart, s);
}

void launch_sym_finalize(const Events &ev, const double *params, int K, const SymGrid &g, bool grad,
                         const double *rowpart, const double *colpart, double *block_out, double *probs,
                         hipStream_t s) {
    dim3 grid((unsigned)((ev.n + kRowsPerBlock - 1) / kRowsPerBlock), (unsigned)K);
    if (grad)
        hipLaunchKernelGGL((sym_finalize_kernel<true>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, rowpart,
                           colpart, block_out, probs);
    else
        hipLaunchKernelGGL((sym_finalize_kernel<false>), grid, dim3(kRowsPerBlock), 0, s, ev, params, g, rowpart,
                           colpart, block_out, probs);
}

} // namespace kern
} // namespace hph
