// Symmetric pair pass: work items, finalize and reduction kernels, host-side plan and launchers.
// The pair-loop device code and its description are in hph_sym_device.h.
#include "hph_sym_device.h"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace hph {
namespace kern {

// Row sums = row-side chunks + diagonal block + column side of every later block. One block of
// 256 rows of one (dataset, vector) per workgroup.
// Moment sources: SPLIT false: all six from the pass's row and column partials. SPLIT true: the
// background moments from bsum and the excitation moments from scache, per-event sums in natural
// units ([k][3 or 1][npad] each); no partials are read.
// sums (may be null): the six moments of dataset 0, vector 0 per event in natural units
// ([6][npad]; without the gradient only moments 0 and 3), for later evaluations that reuse them.
template <bool GRAD, bool SPLIT>
__global__ __launch_bounds__(kRowsPerBlock) void sym_finalize_kernel(
    SymLaunch L, const double *__restrict__ params, const double *__restrict__ rowpart,
    const double *__restrict__ colpart, const double *__restrict__ bsum, const double *__restrict__ scache,
    double *__restrict__ sums, double *__restrict__ block_out, double *__restrict__ probs,
    double *__restrict__ inv_lam) {
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    constexpr int QS = GRAD ? 3 : 1;
    __shared__ double2 sh[kResultWidth * kRowsPerBlock];
    __shared__ int qlist[kQChunk], rlist[kQChunk];
    __shared__ int qcount, rcount;
    const int4 it = L.fin_item[blockIdx.x];
    const SymData &ds = L.data[it.x];
    const int k = it.y, fb = it.z;
    const int i = fb * kRowsPerBlock + threadIdx.x;
    const int n = ds.n;
    const int R = L.R;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // column side: the workgroup's 256 rows are column tile fb of every later block; the later
    // blocks that wrote it are listed kQChunk at a time (the first list here, added below with
    // the row partials, the rest after)
    const int Pw = fb * kRowsPerBlock / R;
    const double *cbase = colpart + ds.col0 + (size_t)k * ds.col_kstride;
    const unsigned char *ftile = L.colflag + ds.flag0 + (size_t)k * ds.flag_kstride + (size_t)fb * ds.nb;
    const int cnt0 = SPLIT ? 0 : flagged_blocks(ftile, Pw + 1, min(ds.nb, Pw + 1 + kQChunk), qlist, &qcount);
    double cs[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) cs[c] = 0.0;
    double a6[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) a6[q] = 0.0;
    if (!SPLIT) {
        // row side: the chunk slots the pass wrote (the block's 256 rows share one row block P)
        const int P = fb * kRowsPerBlock / R;
        const int nch = (P * R + ds.L - 1) / ds.L;
        const unsigned char *rf = L.colflag + ds.rflag0 + (size_t)k * ds.rflag_kstride + (size_t)P * ds.max_chunks;
        add_row_slots<NR, NR>(rowpart + ds.row0 + (size_t)k * ds.row_kstride, rf, nch, ds.max_chunks,
                              (size_t)ds.nb * R, i, i < n, a6, rlist, &rcount);
    }
    if (i < n) {
        if (!SPLIT) {
            add_listed_columns<NC>(cbase, R, i, qlist, cnt0, cs);
        } else {
            // per-event sums, already complete
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                a6[q] = bsum[((L.bsum_shared ? (size_t)0 : (size_t)k * QS) + q) * ds.npad + i];
                a6[NR - QS + q] = scache[((size_t)k * QS + q) * ds.npad + i];
            }
        }
    }
    if (!SPLIT) add_remaining_columns<NC>(cbase, ftile, R, Pw, ds.nb, i, i < n, cs, qlist, &qcount);
    if (i < n) {
        double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (GRAD) {
            s[0] = a6[0] + cs[0];
            s[1] = a6[1] + cs[NC > 1 ? 1 : 0];
            s[2] = a6[2] + cs[NC > 2 ? 2 : 0];
            s[3] = a6[3];
            s[4] = a6[4];
            s[5] = a6[5];
            if (ds.fast && !SPLIT) {
                // the base-2 pass summed e y (y = r'^2 + dt'^2), e dt'^2, e r'^2 and e dt with
                // r' = sx r, dt' = st dt (the excitation's dt is natural)
                const double *pk = params + 6 * (size_t)(ds.param0 + k);
                const double x2 = pk[1] * pk[1], t2 = pk[2] * pk[2];
                const double unit = L.b2tab ? 1.0 / kern::kExp2TabSize : 1.0;
                s[1] = from_base2_sq(s[1] - s[2], x2, unit);
                s[2] = from_base2_sq(s[2], t2, unit);
                s[4] = from_base2_sq(s[4], x2, unit);
            }
        } else {
            s[0] = a6[0] + cs[0];
            s[3] = a6[NR - 1];
        }
        if (sums != nullptr && it.x == 0 && k == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (GRAD || q == 0 || q == 3) sums[(size_t)q * ds.npad + i] = s[q];
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
    size_t row0 = 0, col0 = 0, blk0 = 0, flag0 = 0;
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
        s.flag0 = (long long)flag0;
        // flags [column tile][block]: tile t of the earlier columns (t < (nb - 1) R / kTile) by row block
        s.flag_kstride = (long long)((size_t)std::max(nb[d] - 1, 1) * (R / kTile) * nb[d]);
        s.blk0 = (long long)blk0;
        s.res0 = res0;
        row0 += (size_t)K[d] * s.row_kstride;
        col0 += (size_t)K[d] * s.col_kstride;
        flag0 += (size_t)K[d] * s.flag_kstride;
        s.rflag_kstride = (long long)((size_t)nb[d] * s.max_chunks);
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
    // row-slot flags after every dataset's column flags
    for (SymData &s : plan.data) {
        s.rflag0 = (long long)flag0;
        flag0 += (size_t)s.K * s.rflag_kstride;
    }
    plan.colflag_bytes = flag0;
    plan.block_partials = blk0;
    plan.results = res0;
    return plan;
}

void launch_sym(const SymLaunch &L, const double *params, bool grad, bool fp32, double *rowpart, double *colpart,
                hipStream_t s) {
    if (L.A != 1 && L.A != 2) throw std::invalid_argument("symmetric pass is built for 1 or 2 rows per lane");
    if (L.R != kSymWaves * 64 * L.A) throw std::invalid_argument("symmetric pass: block rows do not match");
    if (L.grid < 1 || L.items < 1) throw std::invalid_argument("symmetric pass: empty launch");
    if (L.fast != 0 && L.b2tab != (fp32 ? 0 : 1))
        throw std::invalid_argument("symmetric pass: base-2 units do not match the precision");
    if (L.pm != kPassBoth) {
        launch_sym_split(L, params, grad, fp32, rowpart, colpart, s);
        return;
    }
    if (L.ktab != 0) {
        if (fp32 || L.fast != 1) throw std::invalid_argument("symmetric pass: the exponent-in-table form is FP64 base-2");
        launch_sym_p<false, 3, kPassBoth>(L, params, grad, rowpart, colpart, s);
        return;
    }
    switch ((fp32 ? 3 : 0) + (L.fast >= 0 && L.fast <= 2 ? L.fast : 0)) {
    case 0: launch_sym_p<false, 0, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    case 1: launch_sym_p<false, 1, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    case 2: launch_sym_p<false, 2, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    case 3: launch_sym_p<true, 0, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    case 4: launch_sym_p<true, 1, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    default: launch_sym_p<true, 2, kPassBoth>(L, params, grad, rowpart, colpart, s); break;
    }
}

template <bool SPLIT>
static void launch_fin(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                       const double *colpart, const double *bsum, const double *scache, double *sums,
                       double *block_out, double *probs, double *inv_lam, hipStream_t s) {
    const dim3 grid((unsigned)L.fin_items), block(kRowsPerBlock);
    if (grad)
        hipLaunchKernelGGL((sym_finalize_kernel<true, SPLIT>), grid, block, 0, s, L, params, rowpart, colpart,
                           bsum, scache, sums, block_out, probs, inv_lam);
    else
        hipLaunchKernelGGL((sym_finalize_kernel<false, SPLIT>), grid, block, 0, s, L, params, rowpart,
                           colpart, bsum, scache, sums, block_out, probs, inv_lam);
}

void launch_sym_finalize(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                         const double *colpart, const double *bsum, const double *scache, double *sums,
                         double *block_out, double *probs, double *inv_lam, hipStream_t s) {
    if ((bsum == nullptr) != (scache == nullptr))
        throw std::invalid_argument("per-event background and excitation sums come together");
    if (bsum == nullptr)
        launch_fin<false>(L, params, grad, rowpart, colpart, bsum, scache, sums, block_out, probs, inv_lam, s);
    else
        launch_fin<true>(L, params, grad, rowpart, colpart, bsum, scache, sums, block_out, probs, inv_lam, s);
}

void launch_sym_reduce(const SymLaunch &L, bool grad, const double *block_out, double *result, hipStream_t s) {
    if (grad)
        hipLaunchKernelGGL((sym_reduce_kernel<kResultWidth>), dim3((unsigned)L.res_items), dim3(kRowsPerBlock), 0, s,
                           L, block_out, result);
    else
        hipLaunchKernelGGL((sym_reduce_kernel<1>), dim3((unsigned)L.res_items), dim3(kRowsPerBlock), 0, s, L,
                           block_out, result);
}

// The work-queue order (SymOrder), one workgroup: per (dataset, vector, block) the two tile
// thresholds, per item its cost, then a counting sort over cost classes. Only a few rounds of
// global loads may stand in its path (the pass waits for it): the searches run on the tiles' last
// times staged in LDS (up to kOrderTiles of them across the datasets; beyond that they read the
// times in device memory), the thresholds stay in LDS (up to kOrderThr (dataset, vector, block)s),
// and with PER > 0 every thread holds its PER items (i = tid + j kOrderThreads) in registers from
// one round of loads to the final scatter; PER = 0 keeps the classes in device memory (plans of
// more than 16 x 1024 items: K-batches, dataset batches). Classes: kOrderClasses linear cost
// classes, each split by the lane that holds the item, so that the 64 lanes of one LDS atomic
// never add to one counter (items of one cost would otherwise serialise it 64-fold).
constexpr int kOrderThreads = 1024;
constexpr int kOrderClasses = 64;
constexpr int kOrderBins = kOrderClasses * 64;
constexpr int kOrderTiles = 4096;
constexpr int kOrderThr = 1920;
template <int PER>
__global__ __launch_bounds__(kOrderThreads) void sym_order_kernel(SymLaunch L, SymOrder O,
                                                                  const double *__restrict__ params) {
    __shared__ int hist[kOrderBins];
    __shared__ double tend[kOrderTiles];
    __shared__ int lthr[2 * kOrderThr];
    __shared__ int wred[kOrderThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int T = kTile, R = L.R, m = L.items;
    // the tables do not overlap: restrict lets the loads of later items go ahead of earlier stores
    const int4 *__restrict__ src = O.src;
    int *__restrict__ gthr = O.thr;
    int *__restrict__ key = O.key;
    int4 *__restrict__ out = O.out;
    // this thread's items, loaded before anything waits on them
    int4 it[PER > 0 ? PER : 1];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = tid + j * kOrderThreads;
        if (i < m) it[j] = src[i];
    }
    // tile-end times: dataset d's tile j (j < n_d / T) at tend[tb_d + j]
    int tiles = 0;
    for (int d = 0; d < O.nd; ++d) tiles += L.data[d].n / T;
    const bool staged = tiles <= kOrderTiles;
    if (staged) {
        int tb = 0;
        for (int d = 0; d < O.nd; ++d) {
            const SymData &sd = L.data[d];
            const int nt = sd.n / T;
            for (int j = tid; j < nt; j += kOrderThreads) tend[tb + j] = sd.t[j * T + T - 1];
            tb += nt;
        }
    }
    for (int b = tid; b < kOrderBins; b += kOrderThreads) hist[b] = 0;
    int *const thr = O.nthr <= kOrderThr ? lthr : gthr;
    __syncthreads();
    for (int q = tid; q < O.nthr; q += kOrderThreads) {
        int d = 0, tb = 0;
        while (d + 1 < O.nd && O.qbase[d + 1] <= q) {
            tb += L.data[d].n / T;
            ++d;
        }
        const SymData &sd = L.data[d];
        const int local = q - O.qbase[d];
        const int k = local / sd.nb, P = local - k * sd.nb;
        const double *pk = params + 6 * (size_t)(sd.param0 + k);
        const double gb = 0.5 * pk[2] * pk[2], gs = pk[3];
        const double zb = sd.zero_arg_b;
        const double xs = excitation_skip_arg(pk, L.dim, sd.n);
        if (P == 0) O.xs[sd.param0 + k] = xs; // one writer per vector
        const double zs = L.cut_s ? xs : kZeroExpArg;
        const double *t = sd.t;
        const double tb0 = t[min(P * R, sd.n - 1)];
        // the block's earlier tiles j < J; their last events j T + T - 1 < P R <= n - 1
        const int J = P * R / T;
        auto tile_end = [&](int j) { return staged ? tend[tb + j] : t[j * T + T - 1]; };
        // the pass's workgroup-wide tests (sym_item), both monotone in j: skipped for j below the
        // first threshold, excitation terms from the second on
        int lo = 0, hi = J;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const double tl = tile_end(mid), gap = tb0 - tl;
            const bool skipped = L.skip_zero && tl < tb0 && gb * gap * gap >= zb && gs >= 0.0 && gs * gap >= zs;
            if (skipped) lo = mid + 1; else hi = mid;
        }
        thr[2 * q] = lo;
        lo = 0;
        hi = J;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const double gap = tb0 - tile_end(mid);
            const bool no_exc = L.skip_zero && gap > 0.0 && gs >= 0.0 && gs * gap >= zs;
            if (no_exc) lo = mid + 1; else hi = mid;
        }
        thr[2 * q + 1] = lo;
    }
    __syncthreads();
    // an item's cost: its chunk's tiles against its block's thresholds (tenths of a tile: 1 per
    // skipped tile, 10 x 10 with background terms only, 10 x 21 with both); a diagonal item
    // 210 rows^2 / (R kTile)
    auto item_cost = [&](int4 v) {
        const SymData &sd = L.data[v.x];
        if (v.z < 0) {
            const double rows = (double)min(R, sd.n - v.y * R);
            return (int)(210.0 * rows * rows / ((double)R * T) + 0.5);
        }
        const int q = O.qbase[v.x] + v.w * sd.nb + v.y;
        const int a = v.z * sd.L / T, b = min((v.z + 1) * sd.L, v.y * R) / T;
        const int j1 = min(max(thr[2 * q], a), b);     // first computed tile
        const int j2 = min(max(thr[2 * q + 1], j1), b); // first with excitation terms
        return (j1 - a) + 100 * (j2 - j1) + 210 * (b - j2);
    };
    int cmax = 0;
    int c[PER > 0 ? PER : 1];
    if (PER > 0) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            c[j] = tid + j * kOrderThreads < m ? item_cost(it[j]) : 0;
            cmax = max(cmax, c[j]);
        }
    } else {
        for (int i = tid; i < m; i += kOrderThreads) {
            const int ci = item_cost(src[i]);
            key[i] = ci;
            cmax = max(cmax, ci);
        }
    }
    for (int o = 32; o > 0; o >>= 1) cmax = max(cmax, __shfl_xor(cmax, o));
    if (lane == 0) wred[w] = cmax;
    __syncthreads();
    cmax = 0;
    for (int v = 0; v < kOrderThreads / 64; ++v) cmax = max(cmax, wred[v]);
    // bin = (class, lane); ascending class = descending cost
    auto bin_of = [&](int ci) {
        const int cls = (int)((long long)ci * (kOrderClasses - 1) / (cmax > 0 ? cmax : 1));
        return (kOrderClasses - 1 - cls) * 64 + lane;
    };
    if (PER > 0) {
        // independent LDS additions, issued back to back
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            c[j] = bin_of(c[j]);
            if (tid + j * kOrderThreads < m) atomicAdd(&hist[c[j]], 1);
        }
    } else {
        for (int i = tid; i < m; i += kOrderThreads) {
            const int b = bin_of(key[i]);
            key[i] = b;
            atomicAdd(&hist[b], 1);
        }
    }
    __syncthreads();
    // exclusive prefix sums of the bins: kOrderBins / kOrderThreads consecutive bins per thread,
    // an inclusive scan across each wave, then across the waves
    {
        constexpr int per = kOrderBins / kOrderThreads;
        int v[per], sum = 0;
#pragma unroll
        for (int b = 0; b < per; ++b) {
            v[b] = hist[tid * per + b];
            sum += v[b];
        }
        int incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        __syncthreads(); // every wave has read the maxima
        if (lane == 63) wred[w] = incl;
        __syncthreads();
        int run = incl - sum;
        for (int v2 = 0; v2 < w; ++v2) run += wred[v2];
#pragma unroll
        for (int b = 0; b < per; ++b) {
            hist[tid * per + b] = run;
            run += v[b];
        }
    }
    __syncthreads();
    if (PER > 0) {
        // every slot taken first (independent LDS atomics, one wait), then the stores
        int slot[PER > 0 ? PER : 1];
#pragma unroll
        for (int j = 0; j < PER; ++j) slot[j] = tid + j * kOrderThreads < m ? atomicAdd(&hist[c[j]], 1) : 0;
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (tid + j * kOrderThreads < m) out[slot[j]] = it[j];
    } else {
        for (int i = tid; i < m; i += kOrderThreads) out[atomicAdd(&hist[key[i]], 1)] = src[i];
    }
}

void launch_sym_order(const SymLaunch &L, const SymOrder &O, const double *params, hipStream_t s) {
    if (L.items < 1 || O.nd < 1 || O.out == nullptr) throw std::invalid_argument("work-queue order: empty plan");
    const int per = (L.items + kOrderThreads - 1) / kOrderThreads;
    const dim3 g(1), b(kOrderThreads);
    if (per <= 4) hipLaunchKernelGGL(sym_order_kernel<4>, g, b, 0, s, L, O, params);
    else if (per <= 8) hipLaunchKernelGGL(sym_order_kernel<8>, g, b, 0, s, L, O, params);
    else if (per <= 12) hipLaunchKernelGGL(sym_order_kernel<12>, g, b, 0, s, L, O, params);
    else if (per <= 16) hipLaunchKernelGGL(sym_order_kernel<16>, g, b, 0, s, L, O, params);
    else hipLaunchKernelGGL(sym_order_kernel<0>, g, b, 0, s, L, O, params);
}

} // namespace kern
} // namespace hph
