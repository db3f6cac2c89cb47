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

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace hph {
namespace kern {


// Per-vector constants of the pair loops, in one of two forms.
//  natural (FAST false): times as given; exponents in e-units with the clamped exp_fp64:
//      background cB0 r^2 + cB1 dt^2 (cB0 = -tauX^2/2, cB1 = -tauT^2/2; times log2 e for the
//      FP32-bulk v_exp_f32), excitation cS0 r^2 + cS1 dt (cS0 = -sigma^2/2, cS1 = -omega).
//  base 2 (FAST true): times arrive as dt' = st dt (time_scale); exponents in 2-units with the
//      clamp-free exp2_fp64: background -(cB0 r^2 + dt'^2) with cB0 = tauX^2 log2(e)/2, excitation
//      cS0 r^2 + cS1 dt' with cS0 = -sigma^2 log2(e)/2 and cS1 = -omega log2(e)/st. The row sums
//      of e dt^2 and e dt then carry st^2 and st, which the finalize kernel divides out.
struct PairConst {
    double cB0, cB1, cS0, cS1;
};

// Background term of one pair and its dt^2 (scaled by st^2 in the base-2 form).
template <bool F32, bool FAST>
__device__ __forceinline__ double background_term(const PairConst &c, double r2, double dt, double &dt2) {
    dt2 = dt * dt;
    if (FAST) {
        const double y = __builtin_fma(c.cB0, r2, dt2);
        if (F32) return (double)exp2_f32((float)(-y));
        return exp2_fp64(-y);
    }
    return background_exp<F32>(c.cB0, c.cB1, r2, dt2);
}

template <bool FAST>
__device__ __forceinline__ double excitation_term(const PairConst &c, double r2, double dt) {
    const double y = __builtin_fma(c.cS0, r2, c.cS1 * dt);
    return FAST ? exp2_fp64(y) : exp_fp64(y);
}

// One 64-column subtile against this lane's A rows, rotating the column accumulators ca.
// DOB: form the background term; SMODE: excitation absent / unmasked / masked; TAIL: the block
// has rows past the last event, whose column contributions are masked by vm; F32: FP32-bulk
// background exponential; FAST: the base-2 form of PairConst.
template <int D, bool GRAD, int A, bool DOB, int SMODE, bool TAIL, bool F32, bool FAST>
__device__ __forceinline__ void sym_subtile(const double *__restrict__ sub, int lane, const double (&xi)[A][D],
                                            const double (&ti)[A], const double (&vm)[A], const PairConst &c,
                                            double (&acc)[A][6], double (&ca)[GRAD ? 3 : 1]) {
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
                double dt2;
                const double eb = background_term<F32, FAST>(c, r2, dt, dt2);
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
                double es = excitation_term<FAST>(c, r2, dt);
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
template <int D, bool GRAD, bool F32, bool FAST>
__device__ __forceinline__ void diag_pass(const double *__restrict__ tile, int cnt, const double (&xi)[D], double ti,
                                          const PairConst &c, double (&acc)[6]) {
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
        double dt2;
        const double eb = background_term<F32, FAST>(c, r2, dt, dt2);
        acc[0] += eb;
        if (GRAD) {
            acc[1] = __builtin_fma(eb, r2, acc[1]);
            acc[2] = __builtin_fma(eb, dt2, acc[2]);
        }
        double es = excitation_term<FAST>(c, r2, dt);
        es = dt > 0.0 ? es : 0.0;
        acc[3] += es;
        if (GRAD) {
            acc[4] = __builtin_fma(es, r2, acc[4]);
            acc[5] = __builtin_fma(es, dt, acc[5]);
        }
    }
}

template <int D, bool GRAD, int A, bool TAIL, bool F32, bool FAST>
__device__ __forceinline__ void sym_dispatch(int mode, const double *__restrict__ sub, int lane,
                                             const double (&xi)[A][D], const double (&ti)[A], const double (&vm)[A],
                                             const PairConst &c, double (&acc)[A][6], double (&ca)[GRAD ? 3 : 1]) {
    switch (mode) {
    case 4 + kSNone:
        sym_subtile<D, GRAD, A, true, kSNone, TAIL, F32, FAST>(sub, lane, xi, ti, vm, c, acc, ca);
        break;
    case 4 + kSFull:
        sym_subtile<D, GRAD, A, true, kSFull, TAIL, F32, FAST>(sub, lane, xi, ti, vm, c, acc, ca);
        break;
    case 4 + kSMixed:
        sym_subtile<D, GRAD, A, true, kSMixed, TAIL, F32, FAST>(sub, lane, xi, ti, vm, c, acc, ca);
        break;
    case kSFull:
        sym_subtile<D, GRAD, A, false, kSFull, TAIL, F32, FAST>(sub, lane, xi, ti, vm, c, acc, ca);
        break;
    case kSMixed:
        sym_subtile<D, GRAD, A, false, kSMixed, TAIL, F32, FAST>(sub, lane, xi, ti, vm, c, acc, ca);
        break;
    default: break; // every term of the subtile is exactly zero
    }
}

// One work item of the symmetric pass: row block P of dataset it.x, vector it.w, against a chunk
// of earlier columns (it.z >= 0) or its own diagonal block (it.z < 0). FAST: the base-2 form.
template <int D, bool GRAD, int A, bool F32, bool FAST>
__device__ __forceinline__ void sym_item(const SymLaunch &L, const int4 it, const double *__restrict__ params,
                                         double *__restrict__ rowpart, double *__restrict__ colpart,
                                         double *__restrict__ tile, double *__restrict__ colbuf, int tid, int lane,
                                         int w) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    constexpr int R = kSymWaves * 64 * A;
    const SymData &ds = L.data[it.x];
    const int P = it.y, chunk = it.z, k = it.w;
    const double *pk = params + 6 * (size_t)(ds.param0 + k);
    const double sigma = pk[0], tauX = pk[1], tauT = pk[2], omega = pk[3];
    const double *__restrict__ ex = ds.x;
    const double *__restrict__ et = ds.t;
    const int n = ds.n, npad = ds.npad;
    const int P0 = P * R;

    // base-2 form: times relative to the block's first event, scaled by st
    const double st = FAST ? time_scale(tauT) : 1.0;
    const double t0 = FAST ? et[min(P0, n - 1)] : 0.0;
    PairConst c;
    if (FAST) {
        c.cB0 = tauX * tauX * (0.5 * kLog2e);
        c.cB1 = 0.0;
        c.cS0 = -sigma * sigma * (0.5 * kLog2e);
        c.cS1 = -omega * kLog2e / st;
    } else {
        c.cB0 = -0.5 * tauX * tauX * (F32 ? kLog2e : 1.0);
        c.cB1 = -0.5 * tauT * tauT * (F32 ? kLog2e : 1.0);
        c.cS0 = -0.5 * sigma * sigma;
        c.cS1 = -omega;
    }
    // skip tests: a tile is skipped when the smallest time gap alone drives every background
    // exponent past the negligible point or every excitation exponent past the underflow
    // point (in the units of the form)
    const double zb = FAST ? ds.zero_arg_b * kLog2e : ds.zero_arg_b;
    const double xs = L.cut_s ? excitation_skip_arg(pk, D, n) : L.zero_arg_s;
    const double zs = FAST ? xs * kLog2e : xs;
    const double gb = FAST ? 1.0 : 0.5 * tauT * tauT; // background exponent per gap^2
    const double gs = FAST ? -c.cS1 : omega;          // excitation exponent per gap
    const int wrow0 = __builtin_amdgcn_readfirstlane(P0 + w * 64 * A);
    double xi[A][D], ti[A], vm[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int r = wrow0 + a * 64 + lane;
        const int rc = min(r, n - 1);
#pragma unroll
        for (int d = 0; d < D; ++d) xi[a][d] = ex[(size_t)d * npad + rc];
        ti[a] = FAST ? (et[rc] - t0) * st : et[rc];
        vm[a] = r < n ? 1.0 : 0.0;
    }
    const bool wave_live = wrow0 < n;
    const double t_lo = FAST ? (et[min(wrow0, n - 1)] - t0) * st : et[min(wrow0, n - 1)];

    double acc[A][6];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[a][q] = 0.0;

    int slot;
    if (chunk < 0) {
        slot = ds.max_chunks;
        const int c1 = min(P0 + R, n);
        for (int j0 = P0; j0 < c1; j0 += kTile) {
            const int cnt = min(kTile, c1 - j0);
            __syncthreads();
            if (tid < cnt) {
#pragma unroll
                for (int d = 0; d < D; ++d) tile[tid * ST + d] = ex[(size_t)d * npad + j0 + tid];
                tile[tid * ST + D] = FAST ? (et[j0 + tid] - t0) * st : et[j0 + tid];
            }
            __syncthreads();
            if (wave_live) {
#pragma unroll
                for (int a = 0; a < A; ++a)
                    diag_pass<D, GRAD, F32, FAST>(tile, cnt, xi[a], ti[a], c, acc[a]);
            }
        }
    } else {
        slot = chunk;
        const int c0 = chunk * ds.L;
        const int c1 = min(c0 + ds.L, P0);
        const bool tail = P0 + R > n;
        double *cp = colpart + ds.col0 + (size_t)k * ds.col_kstride + (size_t)NC * R * ((size_t)P * (P - 1) / 2);
        // the block's first row: the smallest time gap to any earlier tile is its gap
        const double t_blk = FAST ? (et[min(P0, n - 1)] - t0) * st : et[min(P0, n - 1)];
        for (int j0 = c0; j0 < c1; j0 += kTile) {
            // c1 - c0 and P0 are multiples of kTile: every tile here is full.
            if (L.skip_zero) {
                // Workgroup-wide skip, decided as the per-wave test below decides it for the
                // first wave (whose gap is the smallest): when neither term survives there, it
                // survives for no row of the block; the tile's column partials are then zero.
                const double tl = FAST ? (et[j0 + kTile - 1] - t0) * st : et[j0 + kTile - 1];
                const double gap = t_blk - tl;
                if (tl < t_blk && gb * gap * gap >= zb && gs >= 0.0 && gs * gap >= zs) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) cp[(size_t)c * P0 + j0 + tid] = 0.0;
                    continue;
                }
            }
            __syncthreads();
            {
                // subtile sb = tid / 64 holds columns j0 + 64 sb .. + 63 as D + 1 arrays of 128
                // doubles (each 64-column run stored twice); lanes then read consecutive doubles.
                const int e0 = (tid >> 6) * (D + 1) * 128 + (tid & 63);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double v = ex[(size_t)d * npad + j0 + tid];
                    tile[e0 + d * 128] = v;
                    tile[e0 + d * 128 + 64] = v;
                }
                const double tv = FAST ? (et[j0 + tid] - t0) * st : et[j0 + tid];
                tile[e0 + D * 128] = tv;
                tile[e0 + D * 128 + 64] = tv;
            }
            __syncthreads();
            if (wave_live) {
                // columns are earlier than the rows: tl <= t_lo; equal times need the mask
                const double tl = tile[(3 * (D + 1) + D) * 128 + 63];
                int smode = tl < t_lo ? kSFull : kSMixed;
                int dob = 1;
                if (L.skip_zero) {
                    const double gap = tl < t_lo ? t_lo - tl : 0.0;
                    if (gb * gap * gap >= zb) dob = 0;
                    if (smode == kSFull && gs >= 0.0 && gs * gap >= zs) smode = kSNone;
                }
                const int mode = __builtin_amdgcn_readfirstlane(dob * 4 + smode);
                for (int sb = 0; sb < 4; ++sb) {
                    double ca[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) ca[c] = 0.0;
                    const double *sub = tile + sb * (D + 1) * 128;
                    if (tail)
                        sym_dispatch<D, GRAD, A, true, F32, FAST>(mode, sub, lane, xi, ti, vm, c, acc, ca);
                    else
                        sym_dispatch<D, GRAD, A, false, F32, FAST>(mode, sub, lane, xi, ti, vm, c, acc, ca);
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
    const size_t rpad = (size_t)ds.nb * R;
    double *rp = rowpart + ds.row0 + (size_t)k * ds.row_kstride + (size_t)slot * NR * rpad;
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

// Persistent pass: every workgroup takes work items from the queue until it is empty. FM = 0:
// every item in the natural form; 1: every item in the base-2 form; 2: each item in the base-2
// form where its dataset allows it (SymData::fast), else natural (both paths in one kernel, at
// some register cost, so it is used only for mixed dataset tables).
// Three waves per SIMD (at most 168 VGPRs): measured slightly faster than the two that the
// unconstrained allocation (up to 228 VGPRs for 2 rows per lane) leaves room for.
template <int D, bool GRAD, int A, bool F32, int FM>
__global__ __launch_bounds__(kSymWaves * 64) __attribute__((amdgpu_waves_per_eu(FM == 2 ? 2 : 3, 3))) void sym_kernel(
    SymLaunch L, const double *__restrict__ params, double *__restrict__ rowpart, double *__restrict__ colpart) {
    constexpr int ST = TileStride<D>::value;
    constexpr int NC = GRAD ? 3 : 1;
    // diagonal items: kTile events x ST doubles; chunk items: 4 subtiles x (D + 1) x 128 doubles
    __shared__ double tile[(2 * kTile * (D + 1) > kTile * ST) ? 2 * kTile * (D + 1) : kTile * ST];
    __shared__ double colbuf[kSymWaves * NC * kTile];
    __shared__ int s_item;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(L.counter, 1);
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(s_item); // uniform: item data in scalar registers
        __syncthreads();
        if (idx >= L.items) break;
        const int4 it = L.item[idx];
        if (FM == 1 || (FM == 2 && L.data[it.x].fast))
            sym_item<D, GRAD, A, F32, true>(L, it, params, rowpart, colpart, tile, colbuf, tid, lane, w);
        else
            sym_item<D, GRAD, A, F32, false>(L, it, params, rowpart, colpart, tile, colbuf, tid, lane, w);
    }
}

// Zeroes the work-queue counter before a pass (a kernel, so the pass can be captured in a graph).
__global__ void queue_reset_kernel(int *counter) { *counter = 0; }

// Row sums = row-side chunks + diagonal block + column side of every later block. One block of
// 256 rows of one (dataset, vector) per workgroup.
template <bool GRAD>
__global__ __launch_bounds__(kRowsPerBlock) void sym_finalize_kernel(SymLaunch L, const double *__restrict__ params,
                                                                     const double *__restrict__ rowpart,
                                                                     const double *__restrict__ colpart,
                                                                     double *__restrict__ block_out,
                                                                     double *__restrict__ probs,
                                                                     double *__restrict__ inv_lam) {
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    __shared__ double2 sh[kResultWidth * kRowsPerBlock];
    const int4 it = L.fin_item[blockIdx.x];
    const SymData &ds = L.data[it.x];
    const int k = it.y, fb = it.z;
    const int i = fb * kRowsPerBlock + threadIdx.x;
    const int n = ds.n;
    const int R = L.R;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const int P = i / R;
        const int nch = (P * R + ds.L - 1) / ds.L;
        const size_t rpad = (size_t)ds.nb * R;
        const double *rbase = rowpart + ds.row0 + (size_t)k * ds.row_kstride;
        double a6[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) a6[q] = 0.0;
        // unrolled so that several slots' loads are in flight (one grid of 256-row blocks is too
        // small to cover HBM latency with one load per thread)
#pragma unroll 4
        for (int c = 0; c < nch; ++c) {
            const double *rp = rbase + (size_t)c * NR * rpad;
#pragma unroll
            for (int q = 0; q < NR; ++q) a6[q] += rp[(size_t)q * rpad + i];
        }
        {
            const double *rp = rbase + (size_t)ds.max_chunks * NR * rpad;
#pragma unroll
            for (int q = 0; q < NR; ++q) a6[q] += rp[(size_t)q * rpad + i];
        }
        double cs[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) cs[c] = 0.0;
        const double *cbase = colpart + ds.col0 + (size_t)k * ds.col_kstride;
#pragma unroll 4
        for (int Q = P + 1; Q < ds.nb; ++Q) {
            const double *cp = cbase + (size_t)NC * R * ((size_t)Q * (Q - 1) / 2);
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
            if (ds.fast) { // the base-2 pass summed e dt'^2 and e dt' with dt' = st dt
                const double st = time_scale(params[6 * (size_t)(ds.param0 + k) + 2]);
                s[2] /= st * st;
                s[5] /= st;
            }
        } else {
            s[0] = a6[0] + cs[0];
            s[3] = a6[NR - 1];
        }
        double lam;
        const double p =
            row_terms<GRAD>(s, params + 6 * (size_t)(ds.param0 + k), L.dim, ds.t[i], ds.t_last, v, lam);
        if (probs != nullptr && it.x == 0 && k == 0) probs[ds.perm[i]] = p;
        if (inv_lam != nullptr && it.x == 0 && k == 0) inv_lam[i] = 1.0 / lam;
    }
    const size_t blk = (size_t)ds.blk0 + (size_t)k * ds.fin_blocks + fb;
    constexpr int W = GRAD ? 7 : 1;
    double2 r[W];
#pragma unroll
    for (int q = 0; q < W; ++q) r[q] = make_double2(v[q], 0.0);
    block_reduce_dd_w<W>(r, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            double *o = block_out + (blk * kResultWidth + q) * 2;
            o[0] = r[q].x;
            o[1] = r[q].y;
        }
    }
}

// One workgroup per (dataset, vector): fixed-order double-double sum of its block partials plus
// the N (D-1) log(1/sqrt(2 pi)) constant (src/NewHawkes.hpp:659).
template <int W>
__global__ __launch_bounds__(kRowsPerBlock) void sym_reduce_kernel(SymLaunch L, const double *__restrict__ block_out,
                                                                   double *__restrict__ result) {
    __shared__ double2 sh[W * kRowsPerBlock];
    const int2 it = L.res_item[blockIdx.x];
    const SymData &ds = L.data[it.x];
    const int k = it.y;
    const size_t blk0 = (size_t)ds.blk0 + (size_t)k * ds.fin_blocks;
    double2 a[W];
#pragma unroll
    for (int q = 0; q < W; ++q) a[q] = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < ds.fin_blocks; b += kRowsPerBlock) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const double *o = block_out + ((blk0 + b) * kResultWidth + q) * 2;
            a[q] = dd_add(a[q], make_double2(o[0], o[1]));
        }
    }
    block_reduce_dd_w<W>(a, sh);
    if (threadIdx.x == 0) {
        const double c = -0.91893853320467274178 * (double)ds.n * (double)(L.dim - 1);
        a[0] = dd_add(a[0], make_double2(c, 0.0));
#pragma unroll
        for (int q = 0; q < W; ++q) result[(size_t)(ds.res0 + k) * kResultWidth + q] = a[q].x + a[q].y;
    }
}

// ---------------- host side ----------------

static long long sym_items_for(int nb, int R, int L) {
    long long items = nb; // diagonal items
    for (int P = 1; P < nb; ++P) items += ((long long)P * R + L - 1) / L;
    return items;
}

// Chunk length per dataset: the longest (multiple of kTile, at most the earlier-column span) that
// still gives at least target_items work items over all datasets and vectors.
SymPlan plan_sym(const std::vector<int> &n, const std::vector<int> &K, int rows_per_lane, int target_items,
                 bool grad) {
    if (rows_per_lane != 1 && rows_per_lane != 2) throw std::invalid_argument("rows per lane must be 1 or 2");
    if (n.size() != K.size() || n.empty()) throw std::invalid_argument("dataset sizes and vector counts differ");
    SymPlan plan;
    plan.A = rows_per_lane;
    plan.R = kSymWaves * 64 * plan.A;
    const int R = plan.R;
    const int nd = (int)n.size();
    std::vector<int> nb(nd), maxlen(nd);
    for (int d = 0; d < nd; ++d) {
        if (n[d] < 1 || K[d] < 1) throw std::invalid_argument("every dataset needs events and a vector");
        nb[d] = (n[d] + R - 1) / R;
        maxlen[d] = ((nb[d] > 1 ? (nb[d] - 1) * R : R) + kTile - 1) / kTile * kTile;
    }
    int L = 0;
    for (int d = 0; d < nd; ++d) L = std::max(L, maxlen[d]);
    auto total_items = [&](int len) {
        long long t = 0;
        for (int d = 0; d < nd; ++d) t += (long long)K[d] * sym_items_for(nb[d], R, std::min(len, maxlen[d]));
        return t;
    };
    while (L > kTile && total_items(L) < target_items) L = ((L / 2 + kTile - 1) / kTile) * kTile;

    const int NR = grad ? 6 : 2, NC = grad ? 3 : 1;
    size_t row0 = 0, col0 = 0, blk0 = 0;
    int res0 = 0, param0 = 0;
    struct Cost {
        long long cost;
        int4 item;
    };
    std::vector<Cost> items;
    for (int d = 0; d < nd; ++d) {
        SymData s{};
        s.n = n[d];
        s.nb = nb[d];
        s.L = std::min(L, maxlen[d]);
        s.max_chunks = nb[d] > 1 ? ((nb[d] - 1) * R + s.L - 1) / s.L : 1;
        s.K = K[d];
        s.param0 = param0;
        s.fin_blocks = (n[d] + kRowsPerBlock - 1) / kRowsPerBlock;
        const size_t rpad = (size_t)nb[d] * R;
        s.row0 = (long long)row0;
        s.row_kstride = (long long)((size_t)(s.max_chunks + 1) * NR * rpad);
        s.col0 = (long long)col0;
        s.col_kstride = (long long)((size_t)NC * R * ((size_t)nb[d] * (nb[d] - 1) / 2));
        s.blk0 = (long long)blk0;
        s.res0 = res0;
        row0 += (size_t)K[d] * s.row_kstride;
        col0 += (size_t)K[d] * s.col_kstride;
        blk0 += (size_t)K[d] * s.fin_blocks;
        res0 += K[d];
        param0 += K[d];
        for (int k = 0; k < K[d]; ++k) {
            for (int P = nb[d] - 1; P >= 0; --P) {
                const int P0 = P * R;
                const int nch = P > 0 ? (P0 + s.L - 1) / s.L : 0;
                for (int c = 0; c < nch; ++c) {
                    const int cols = std::min(s.L, P0 - c * s.L);
                    items.push_back({(long long)R * cols, make_int4(d, P, c, k)});
                }
                const int rows = std::min(R, n[d] - P0);
                items.push_back({(long long)rows * rows, make_int4(d, P, -1, k)});
            }
            for (int fb = 0; fb < s.fin_blocks; ++fb) plan.fin_items.push_back(make_int4(d, k, fb, 0));
            plan.res_items.push_back(make_int2(d, k));
        }
        plan.data.push_back(s);
    }
    std::stable_sort(items.begin(), items.end(), [](const Cost &a, const Cost &b) { return a.cost > b.cost; });
    plan.items.reserve(items.size());
    for (const Cost &c : items) plan.items.push_back(c.item);
    plan.rowpart_doubles = row0;
    plan.colpart_doubles = col0;
    plan.block_partials = blk0;
    plan.results = res0;
    return plan;
}

template <int D, bool GRAD, bool F32, int FM>
static void launch_sym_d(const SymLaunch &L, const double *params, double *rowpart, double *colpart, hipStream_t s) {
    if (L.A == 2)
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 2, F32, FM>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s, L,
                           params, rowpart, colpart);
    else
        hipLaunchKernelGGL((sym_kernel<D, GRAD, 1, F32, FM>), dim3((unsigned)L.grid), dim3(kSymWaves * 64), 0, s, L,
                           params, rowpart, colpart);
}

template <bool F32, int FM>
static void launch_sym_p(const SymLaunch &L, const double *params, bool grad, double *rowpart, double *colpart,
                         hipStream_t s) {
    switch (L.dim * 2 + (grad ? 1 : 0)) {
    case 2: launch_sym_d<1, false, F32, FM>(L, params, rowpart, colpart, s); break;
    case 3: launch_sym_d<1, true, F32, FM>(L, params, rowpart, colpart, s); break;
    case 4: launch_sym_d<2, false, F32, FM>(L, params, rowpart, colpart, s); break;
    case 5: launch_sym_d<2, true, F32, FM>(L, params, rowpart, colpart, s); break;
    case 6: launch_sym_d<3, false, F32, FM>(L, params, rowpart, colpart, s); break;
    case 7: launch_sym_d<3, true, F32, FM>(L, params, rowpart, colpart, s); break;
    default: throw std::invalid_argument("symmetric pass is built for dimensions 1..3");
    }
}

void launch_queue_reset(const SymLaunch &L, hipStream_t s) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, s, L.counter);
}

void launch_sym(const SymLaunch &L, const double *params, bool grad, bool fp32, double *rowpart, double *colpart,
                hipStream_t s) {
    if (L.A != 1 && L.A != 2) throw std::invalid_argument("symmetric pass is built for 1 or 2 rows per lane");
    if (L.R != kSymWaves * 64 * L.A) throw std::invalid_argument("symmetric pass: block rows do not match");
    if (L.grid < 1 || L.items < 1) throw std::invalid_argument("symmetric pass: empty launch");
    launch_queue_reset(L, s);
    switch ((fp32 ? 3 : 0) + (L.fast >= 0 && L.fast <= 2 ? L.fast : 0)) {
    case 0: launch_sym_p<false, 0>(L, params, grad, rowpart, colpart, s); break;
    case 1: launch_sym_p<false, 1>(L, params, grad, rowpart, colpart, s); break;
    case 2: launch_sym_p<false, 2>(L, params, grad, rowpart, colpart, s); break;
    case 3: launch_sym_p<true, 0>(L, params, grad, rowpart, colpart, s); break;
    case 4: launch_sym_p<true, 1>(L, params, grad, rowpart, colpart, s); break;
    default: launch_sym_p<true, 2>(L, params, grad, rowpart, colpart, s); break;
    }
}

void launch_sym_finalize(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                         const double *colpart, double *block_out, double *probs, double *inv_lam, hipStream_t s) {
    if (grad)
        hipLaunchKernelGGL((sym_finalize_kernel<true>), dim3((unsigned)L.fin_items), dim3(kRowsPerBlock), 0, s, L,
                           params, rowpart, colpart, block_out, probs, inv_lam);
    else
        hipLaunchKernelGGL((sym_finalize_kernel<false>), dim3((unsigned)L.fin_items), dim3(kRowsPerBlock), 0, s, L,
                           params, rowpart, colpart, block_out, probs, inv_lam);
}

void launch_sym_reduce(const SymLaunch &L, bool grad, const double *block_out, double *result, hipStream_t s) {
    if (grad)
        hipLaunchKernelGGL((sym_reduce_kernel<kResultWidth>), dim3((unsigned)L.res_items), dim3(kRowsPerBlock), 0, s,
                           L, block_out, result);
    else
        hipLaunchKernelGGL((sym_reduce_kernel<1>), dim3((unsigned)L.res_items), dim3(kRowsPerBlock), 0, s, L,
                           block_out, result);
}

} // namespace kern
} // namespace hph
