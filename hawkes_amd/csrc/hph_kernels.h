// Launch interface between the engine (hph_engine.cpp) and the HIP kernels (hph_kernels.hip).
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>

#include <vector>

namespace hph {
namespace kern {

constexpr int kRowsPerBlock = 256; // one event row per thread, 4 wavefronts of 64
constexpr int kTile = 256;         // events per LDS tile
constexpr int kMaxDim = 8;
constexpr int kResultWidth = 7;    // loglik + 6 gradient components

// Event data of one dataset on the device, time-sorted, structure of arrays.
struct Events {
    const double *x;   // D planes of npad doubles: x[d * npad + i]
    const double *t;   // npad doubles
    const int *perm;   // sorted position -> caller's index
    int n;             // events
    int npad;          // n rounded up to kRowsPerBlock
    int dim;           // D
    double t_last;     // the caller's last time (T of the model)
    const double *t_last_at = nullptr; // device copy of T, read instead of t_last when set
};

// Work decomposition of the pair pass.
struct PairGrid {
    int row_blocks;    // ceil(n / kRowsPerBlock) (finalize blocks)
    int chunks;        // column chunks per row block
    int chunk_len;     // events per chunk (multiple of kTile; kNarrow in the narrow layout)
    int narrow;        // 1: work items of kNarrow rows x kTile columns, 4 wavefronts of 64 columns
    int skip_zero;     // 1: skip tiles whose terms are all negligible (background) or zero (excitation)
    double zero_arg_b; // background tiles whose every exponent is <= -zero_arg_b are skipped
    double zero_arg_s; // excitation exp arguments <= -zero_arg_s give exactly +0.0
    int cut_s;         // 1: excitation tiles below the negligible point are skipped as well
};

// exp(x) is exactly +0.0 in FP64 for x <= -745.14 (both here and in the reference's exp);
// an excitation tile is skipped only when every one of its arguments is <= -kZeroExpArg.
constexpr double kZeroExpArg = 746.0;
// FP32-bulk mode (background term): v_exp_f32 (2^y) is exactly +0.0 for y <= -150, i.e. x <= -103.98.
constexpr double kZeroExpArg32 = 105.0;

// Background tiles are skipped where every term is below e^-X, X = min(exact-zero point,
// 45 + ln n). Bound: a skipped pair has b = e^-(a + c) with a = tauX^2 r^2/2 >= 0 and
// c = tauT^2 dt^2/2 >= X, so b, b 2a and b 2c (the pair's terms in sum b, tauX^2 sum b r^2 and
// tauT^2 sum b dt^2) are each at most 2X e^-X. Every event's own pair is in its sums (b = 1), so
// the n pairs of one event change those sums by at most n 2X e^-X = 2X e^-45 < 4e-18 relative
// to sum b (X <= 67 for n <= 2^31): far below one FP64 rounding (1.1e-16), and still ~3e-15 where
// a gradient component cancels to 1e-3 of its summation scale. The excitation term has no such floor and
// is skipped only where it is exactly zero.
inline double background_skip_arg(long long n, bool fp32) {
    const double exact = fp32 ? kZeroExpArg32 : kZeroExpArg;
    const double negl = 45.0 + std::log((double)(n > 1 ? n : 1));
    return negl < exact ? negl : exact;
}

// Excitation tiles may be skipped where every term is below e^-X_S, X_S = min(exact-zero point,
// 50 + ln n + max(0, ln(|cS| / |cB|))) with cB = mu0 tauX^D tauT / (2 pi) and
// cS = theta omega sigma^D / sqrt(2 pi) (the constants of lambda = cB sum eB + cS sum eS). Bound:
// every event's own background pair gives lambda_i >= |cB|, so the skipped terms (each <= e^-X_S;
// their e dt and e r^2 moments are likewise bounded, omega dt >= X_S) change lambda_i, and the
// excitation parts of the gradient terms, by at most |cS| n e^-X_S <= e^-50 |cB| = 2e-22 |cB|,
// i.e. relatively by under 2e-22. The self-excitation probability S/(B+S) then moves by under
// 2e-22 in absolute terms, which is not a relative bound for tiny probabilities: passes that
// return probabilities skip exact zeros only, or stop each event's excitation sum relative to
// itself (launch_excitation_band, cut 3). Without a background (cB = 0) nothing is cut, and
// neither when cB and cS differ in sign (theta, omega, mu0 or a precision negative: the rate is
// then a difference and the floor lambda_i >= |cB| does not hold).
__host__ __device__ inline double excitation_skip_arg(const double *pk, int D, long long n) {
    double sD = 1.0, xD = 1.0;
    for (int q = 0; q < D; ++q) {
        sD *= pk[0];
        xD *= pk[1];
    }
    const double cBs = pk[5] * xD * pk[2], cSs = pk[4] * pk[3] * sD;
    if (cSs != 0.0 && (cBs > 0.0) != (cSs > 0.0)) return kZeroExpArg;
    const double cB = fabs(cBs) * 0.15915494309189533577;
    const double cS = fabs(cSs) * 0.39894228040143267794;
    if (!(cB > 0.0) || !(cS < 1e300)) return kZeroExpArg;
    const double r = cS > 0.0 ? log(cS / cB) : 0.0;
    const double x = 50.0 + log((double)(n > 1 ? n : 1)) + (r > 0.0 ? r : 0.0);
    return x < kZeroExpArg ? x : kZeroExpArg;
}

// Narrow layout: for small N the 256 x 256 work items cannot fill the device (N = 1000 gives 16
// of them, each lane looping over 256 columns); items of 64 rows x 256 columns whose four
// wavefronts take 64 columns each give 4x the items and a quarter of the per-lane loop.
constexpr int kNarrow = 64;

PairGrid plan_pair_grid(int n, int K, int target_blocks);

// Pair pass: per (k, chunk, row) sums of exp terms and their moments.
//   acc 0: sum exp(-(tauX^2 r^2 + tauT^2 dt^2)/2)      (acc 1: * r^2, acc 2: * dt^2)
//   acc 3: sum_{dt>0} exp(-omega dt - sigma^2 r^2/2)   (acc 4: * r^2, acc 5: * dt)
// Without the gradient only acc 0 and acc 3 are formed (stored as slots 0 and 1).
// part layout: [K][chunks][nacc][npad].
void launch_pair(const Events &ev, const double *params, int K, const PairGrid &g, bool grad, bool fp32,
                 double *part, hipStream_t s);

// Per-row log-likelihood terms and gradient terms, reduced per block in double-double.
// block_out layout: [K][row_blocks][7][2]. probs (may be null) receives p_i for k == 0 in the
// caller's order, inv_lam (may be null) 1 / lambda_i for k == 0 in sorted order, sums (may be
// null) the six moments per event of k == 0 ([6][npad], natural units; without the gradient
// moments 0 and 3).
void launch_finalize(const Events &ev, const double *params, int K, const PairGrid &g, bool grad,
                     const double *part, double *block_out, double *probs, double *inv_lam, double *sums,
                     hipStream_t s);

// Sums the per-block partials (nblocks finalize blocks of 256 rows) per k in a fixed order:
// result layout [K][7] (only [k][0] without grad).
void launch_reduce(const Events &ev, int K, int nblocks, bool grad, const double *block_out, double *result,
                   hipStream_t s);

// ---- Location gradient (hph_locgrad.hip) ----
// Excitation sums per event of K vectors (dimensions 1..3), natural units, out [K][3 or 1][npad]
// (S, S r^2, S dt; without the gradient S only) from the earlier events of each 64-event group,
// per-pair dt > 0 mask. cut 0: within the likelihood passes' excitation cut; 1: within the
// exact-zero point; 2: every earlier event; 3 (probability requests, sums without moments): within
// the exact-zero point, and each group of 64 events stops as soon as, for every one of its events,
// the earlier events not yet added could change its sum by at most 2^-60 of it (a relative bound
// per event; used where theta omega sigma^D > 0 and mu0 tauX^D tauT >= 0, else as cut 1).
void launch_excitation_band(const Events &ev, const double *params, int K, bool grad, int cut, double *out,
                            hipStream_t s);

// w[i] = 1 / lambda_i (sorted order, npad entries) from a rate pass (launch_pair without
// gradient, K = 1, grid g) whose partials are rate_part.
void launch_inv_rate(const Events &ev, const double *params, const PairGrid &g, const double *rate_part, double *w,
                     hipStream_t s);
// d ll / d x into out (N x D, caller's order); part: chunk partials, g.chunks x D x npad doubles.
void launch_locgrad(const Events &ev, const double *params, const double *w, const PairGrid &g, double *part,
                    double *out, hipStream_t s);

// ---- Symmetric pass (hph_sym.hip): background exps once per unordered pair ----
//
// The pass runs over a table of datasets (one for an engine, many for a batch); every dataset
// has its own events, parameter vectors and slices of the partial buffers. Work items
// (dataset, row block, chunk of earlier columns | -1 for the diagonal block, parameter vector)
// are taken from a queue by a persistent grid, heaviest first: the plan orders them by nominal
// size, and launch_sym_order reorders them on the device by their expected cost under the call's
// parameters before each pass in time order (the culled pass's order is formed on the host,
// SymRunner::order_queue_culled). Any order gives the same results.

// One dataset of a symmetric pass (device table entry).
struct SymData {
    const double *x;        // D planes of npad doubles (time-sorted)
    const double *t;        // npad doubles (sorted)
    const int *perm;        // sorted position -> caller's index
    int n, npad;            // events; padded length of x and t
    int nb;                 // row blocks of R rows
    int L;                  // earlier columns per chunk item (multiple of kTile)
    int max_chunks;         // row-partial slots; the diagonal block uses slot max_chunks
    int K;                  // parameter vectors of this dataset
    int param0;             // index of its first parameter vector
    int fin_blocks;         // ceil(n / 256)
    double t_last;          // the caller's last time (T of the model)
    double zero_arg_b;      // background skip point of this dataset (background_skip_arg)
    int fast;               // 1: this dataset's items run the base-2 form (SymLaunch::fast set)
    const double *gstat;    // background pass on the culled order (SymLaunch::pm = kPassCulledB):
                            // per group of 64 events, the minima of the D coordinates and of time,
                            // then their maxima (2 D + 2 doubles); perm then maps a position of
                            // the culled order to the time-sorted position
    long long row0, row_kstride; // row partials: doubles offset of k = 0 and stride per k
    long long col0, col_kstride; // column partials: the same
    long long flag0, flag_kstride; // column-tile flags (bytes): the same
    long long rflag0, rflag_kstride; // row-slot flags (bytes, in the same buffer): per (k, block P,
                                     // chunk c) at rflag0 + k rflag_kstride + P max_chunks + c
    long long blk0;         // index of its (k = 0, block 0) block partial
    int res0;               // index of its (k = 0) result
};

struct SymLaunch {
    const SymData *data;    // device table
    const int4 *item;       // (dataset, block, chunk | -1, k), heaviest first (any order is correct)
    int items;
    int *counter;           // work-queue counter, zeroed by launch_sym before the pass
    int grid;               // persistent workgroups launched
    int dim, A, R;
    int skip_zero;          // as PairGrid::skip_zero (background point: SymData::zero_arg_b)
    double zero_arg_s;      // as PairGrid::zero_arg_s
    int cut_s;              // as PairGrid::cut_s
    int fast;               // base-2 form (scaled times, clamp-free exp2; the host has checked the
                            // exponents, order and spans): 0 no dataset, 1 every dataset, 2 some
                            // (per SymData::fast)
    const int4 *fin_item;   // (dataset, k, finalize block, 0)
    int fin_items;
    const int2 *res_item;   // (dataset, k)
    int res_items;
    int pm;                 // pass mode: kPassBoth or kPassCulledB
    unsigned char *colflag; // per (earlier column tile t, block P), at flag0 + k flag_kstride +
                            // t nb + P: 1 when the pass wrote the tile's column partials, 0 when it
                            // skipped the tile (its partials are then zero and left unwritten);
                            // readers skip unflagged tiles. After them, the row-slot flags
                            // (SymData::rflag0): 2 when a chunk item wrote all its row partials,
                            // 1 when it wrote the background moments only (it formed no
                            // excitation term: those moments are zero and unwritten), 0 when the
                            // pass skipped every tile of the item (partials unwritten)
    int bsum_shared;        // finalize from per-event sums: 1 when bsum holds one set of background
                            // moments for every vector (vectors sharing tauX and tauT), 0 one per vector
    int b2tab;              // 1: the base-2 form ran in FP64 with exponents scaled by 256 (the table
                            // exponential exp2_neg_tab); its squared moments are 256 times the
                            // unscaled ones. 0: FP32-bulk (unscaled). Set from the precision by the
                            // host; launch_sym checks that it matches.
    const double *xs;       // per parameter vector (index SymData::param0 + k): its excitation skip
                            // point excitation_skip_arg, written by launch_sym_order before a
                            // pass in time order (read where cut_s is set)
    int ktab;               // 1: every background exponent of every pair of the launch lies in
                            // exp2_neg_ktab's range (host bound from the data's extents and each
                            // vector's tauX, tauT; FP64 base-2 form, time-sorted pass only): the
                            // pass runs exp2_neg_ktab (bitwise exp2_neg_tab's results)
};

// Pass modes of the symmetric pass. kPassBoth: background and excitation terms of the time-sorted
// events in one pass. For spatially extended data the two terms run as separate passes:
// kPassCulledB forms the background terms on the culled order (the time-sorted events cut into
// windows of consecutive events, each window ordered along a Morton curve of its locations, so
// that groups of 64 events are compact in space and time) and skips every (wave, 64-column group)
// and (block, tile) pair whose bounding boxes put every background exponent past the negligible
// point; the excitation terms then come from launch_excitation_band over the time-sorted events.
constexpr int kPassBoth = 0;
constexpr int kPassCulledB = 1;
constexpr int kGroup = 64; // events per bounding-box group of the culled order
// Box table of the culled order (SymData::gstat): one box per group of kGroup events for
// culled_groups(n) groups (every group of the 512-event row blocks), then one per tile of kTile
// events (culled_groups(n) kGroup / kTile of them), then one per supertile of kSuperTile events
// (the last one partial), so that the background pass tests a far supertile once.
constexpr int kSuperTile = 4096;
__host__ __device__ inline int culled_groups(int n) { return (n + 511) / 512 * (512 / kGroup); }

// Host-side plan of a symmetric pass over datasets of sizes n[d] with K[d] vectors each.
struct SymPlan {
    int A = 2, R = 512;
    std::vector<SymData> data;   // pointers filled in by the caller
    std::vector<int4> items, fin_items;
    std::vector<int2> res_items;
    size_t rowpart_doubles = 0, colpart_doubles = 0, block_partials = 0, colflag_bytes = 0;
    int results = 0;
};

SymPlan plan_sym(const std::vector<int> &n, const std::vector<int> &K, int rows_per_lane, int target_items,
                 bool grad);
void launch_sym(const SymLaunch &L, const double *params, bool grad, bool fp32, double *rowpart, double *colpart,
                hipStream_t s);
// Finalize (per row, then per 256-row block in double-double) and reduction (per dataset and
// vector: result[(res0 + k) * 7 + q]); probs and inv_lam (may be null) receive dataset 0,
// vector 0 (as launch_finalize).
// bsum, scache (both null, or both set): per-event background and excitation sums in natural units
// ([k][3 or 1][npad]; launch_sym_bsums / launch_excitation_band, or the sums of an earlier
// evaluation; bsum without the [k] index when L.bsum_shared); when set, the finalize takes the
// moments from them instead of the pass's partials.
// sums (may be null): receives the six moments per event of dataset 0, vector 0 in natural units
// ([6][npad]; without the gradient moments 0 and 3).
void launch_sym_finalize(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                         const double *colpart, const double *bsum, const double *scache, double *sums,
                         double *block_out, double *probs, double *inv_lam, hipStream_t s);
// Background row sums of a kPassCulledB pass over dataset 0 (row partials + column partials),
// scattered to time-sorted positions in natural units: bsum[(k * NB + q) * npad + sorted index],
// NB = 3 with the gradient (sum, r^2 and dt^2 moments), else 1.
void launch_sym_bsums(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                      const double *colpart, double *bsum, hipStream_t s);
void launch_sym_reduce(const SymLaunch &L, bool grad, const double *block_out, double *result, hipStream_t s);

// Work-queue order of a time-sorted pass for the call's parameters, formed on the device before
// the pass: each item's expected cost counts its tiles that pass the pass's own workgroup-wide
// skip tests (1 per skipped tile, 100 per tile with background terms only, 210 per tile with both
// terms; a diagonal item's cost is fixed by its rows), found by binary search on the block's first
// time (the tests are monotone in the tile index), and the items are counting-sorted by cost,
// heaviest first, into L.item. Items of one cost class may land in any order: every item writes
// fixed locations, so the results do not depend on the order (only the pass's tail does). It also
// writes each parameter vector's excitation skip point (O.xs), which the pass then reads instead
// of evaluating its logarithms per work item.
struct SymOrder {
    const int4 *src;  // the plan's items, in plan order
    const int *qbase; // per dataset: its first (vector, block) index (nd + 1 entries, the last = nthr)
    int *thr;         // scratch, 2 ints per (dataset, vector, block): the first tile not skipped and
                      // the first with excitation terms (used when they do not fit in LDS)
    int *key;         // scratch, 1 int per item (plans of more than 16 x 1024 items)
    int4 *out;        // the queue the pass reads (L.item)
    double *xs;       // per parameter vector: its excitation skip point (the pass's L.xs)
    int nd, nthr;
};
void launch_sym_order(const SymLaunch &L, const SymOrder &O, const double *params, hipStream_t s);
// The work-queue counter (SymLaunch::counter) is zero before every persistent pass and is
// returned to zero by the pass itself (queue_next, hph_sym_device.h): no reset between launches.
// Symmetric location-gradient pass over dataset 0, vector 0 of a plan made with grad = true
// (D <= 3): w = 1/lambda in sorted order; out = d ll / d x, N x D in the caller's order.
void launch_sym_locgrad(const SymLaunch &L, const double *params, const double *w, double *rowpart,
                        double *colpart, double *out, hipStream_t s);

} // namespace kern

namespace noalloc {
// hph_log_likelihood_noalloc (hph_noalloc.hip): K vectors of one dataset from device-resident
// inputs, every intermediate in the library's static scratch array; throws std::invalid_argument
// for arguments and sizes the array cannot hold, std::runtime_error for device failures.
void evaluate(int D, int N, const double *locations, const double *times, const double *params, int K, long flags,
              double *result, bool grad, hipStream_t stream);
} // namespace noalloc
} // namespace hph
