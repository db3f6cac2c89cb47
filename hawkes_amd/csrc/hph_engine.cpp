// HipHawkes: the MI355X engine behind the AbstractHawkes interface, the factory, and the C ABI.
//
// Interface and semantics follow src/AbstractHawkes.hpp:32-64 and src/NewHawkes.hpp:244-339:
//   * updateLocations(-1, all, N*D) or (i, one, D), row-major (src/NewHawkes.hpp:244-276);
//   * setTimesData(N values), T = last element (src/NewHawkes.hpp:320-323, :653);
//   * setParameters(6 values) in the order (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0)
//     (src/NewHawkes.hpp:331-339);
//   * getProbsSelfExcite writes N values in the caller's order (src/NewHawkes.hpp:325-329).
// Internally the events are kept time-sorted on the device (structure of arrays); the
// log-likelihood is a sum over rows and so is unchanged by the order, and probabilities are
// scattered back to the caller's order by the finalize kernel.
#include "hph.h"
#include "hph_engine.hpp"
#include "hph_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace hph {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceOutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NotInitialized : std::logic_error {
    using std::logic_error::logic_error;
};

static void check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) throw DeviceOutOfMemory(msg);
    throw DeviceError(msg);
}

// Device buffer that grows on demand and never shrinks.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t want) {
        if (want <= bytes) return;
        if (ptr) check(hipFree(ptr), "hipFree");
        ptr = nullptr;
        bytes = 0;
        check(hipMalloc(&ptr, want), "hipMalloc");
        bytes = want;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host buffer that grows on demand.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t want) {
        if (want <= bytes) return;
        if (ptr) check(hipHostFree(ptr), "hipHostFree");
        ptr = nullptr;
        bytes = 0;
        check(hipHostMalloc(&ptr, want, hipHostMallocDefault), "hipHostMalloc");
        bytes = want;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Size bounds of a dataset's exponents (squared diagonal of the locations' bounding box, time
// span) and whether its distinct times are far enough apart for the base-2 pair form. With them
// the host picks the pass variant for given parameters.
struct Extents {
    double r2max = 0.0, span = 0.0;
    double block_span = 0.0;  // largest time span of 512 consecutive sorted events
    double window_span = 0.0; // largest time span of a window of the split passes' culled order
    bool gaps_ok = true;      // consecutive distinct times differ by at least 8 ulp(max |t|)
};

// Events per window of the culled order: N / 8 rounded down to a power of two, in
// [1024, 32768] (longer windows make groups smaller in space but longer in time).
static int cull_window(int n) {
    int w = 1024;
    while (w < 32768 && 2 * w <= n / 8) w *= 2;
    return w;
}

// locs: N x D in any order; sorted_t: the N times in ascending order.
static Extents extents_of(const std::vector<double> &locs, const double *sorted_t, size_t n, int D) {
    Extents e;
    if (n == 0) return e;
    for (int d = 0; d < D; ++d) {
        double lo = locs[d], hi = locs[d];
        for (size_t i = 1; i < n; ++i) {
            lo = std::min(lo, locs[i * D + d]);
            hi = std::max(hi, locs[i * D + d]);
        }
        e.r2max += (hi - lo) * (hi - lo);
    }
    e.span = sorted_t[n - 1] - sorted_t[0];
    const double tabs = std::max(std::fabs(sorted_t[0]), std::fabs(sorted_t[n - 1]));
    const double ulp = std::nextafter(tabs, HUGE_VAL) - tabs;
    for (size_t i = 1; i < n && e.gaps_ok; ++i) {
        const double g = sorted_t[i] - sorted_t[i - 1];
        if (g > 0.0 && g < 8.0 * ulp) e.gaps_ok = false;
    }
    for (size_t b = 0; b < n; b += 512) e.block_span = std::max(e.block_span, sorted_t[std::min(b + 511, n - 1)] - sorted_t[b]);
    const size_t W = (size_t)cull_window((int)n);
    for (size_t b = 0; b < n; b += W) e.window_span = std::max(e.window_span, sorted_t[std::min(b + W - 1, n - 1)] - sorted_t[b]);
    return e;
}

// True when the pass may use the base-2 form (scaled coordinates and times, clamp-free exp2):
// every background (tauX^2 r^2/2 + tauT^2 dt^2/2) exponent of the K vectors is below 5.5e6 and
// every excitation (|omega| dt + sigma^2 r^2/2) exponent below 1e9 in size, the scales st = |tauT| sqrt(log2(e)/2) and
// sx = |tauX| sqrt(log2(e)/2) lie in [1e-100, 1e100] with omega log2(e) / st and sigma^2 / tauX^2
// finite, and distinct times are at least 8 ulp(max |t|) apart: then (t - t0) st, rounded, keeps
// every strict order and every tie of the times, so the dt > 0 mask of the excitation term is
// decided as on the original times. Accuracy: the scaled times carry rounding of order
// eps |t - t0| st, t0 = the block's first event, so an excitation exponent omega dt is off by up
// to ~2 eps log2(e) |omega| (block time span); the form is used only while that stays below
// 1e-13 (|omega| x span <= 300), i.e. excitation terms within 1e-13 relative. Likewise the
// scaled coordinates (x - x0) sx carry rounding of order eps |x - x0| sx, x0 = the block's first
// event: a term whose spatial exponent is of order one is off by ~eps x (scale x extent) relative,
// kept below 1e-13 by sx x diagonal <= 300 and sigma sqrt(log2(e)/2) x diagonal <= 300.
// The split passes' background pass runs on the culled order, whose row blocks are ordered along
// a space-filling curve inside a window of consecutive sorted events: their times, and so t - t0,
// span the whole window, not 512 events. split adds the condition st x (window time span) <= 300,
// which keeps the scaled times' rounding as small as sx x diagonal <= 300 keeps the coordinates'.
// The table exponential takes exponents of one sign: the excitation exponent
// -(omega dt + sigma^2 r^2 / 2) of a causal pair is <= 0 only for omega >= 0, so omega < 0
// (accepted by setParameters, unphysical) takes the natural form.
static bool fast_ok(const Extents &e, const double *params, int K, bool split = false) {
    if (!e.gaps_ok) return false;
    const double diag = std::sqrt(e.r2max);
    for (int k = 0; k < K; ++k) {
        const double *p = params + 6 * (size_t)k;
        if (!(p[3] >= 0.0)) return false;
        if (!(std::fabs(p[3]) * e.block_span <= 300.0)) return false;
        const double aB = 0.5 * p[1] * p[1] * e.r2max + 0.5 * p[2] * p[2] * e.span * e.span;
        const double aS = std::fabs(p[3]) * e.span + 0.5 * p[0] * p[0] * e.r2max;
        // background: 256 log2(e) aB < 2^31 (the FP64 table exponential converts 256 x its base-2
        // exponent to an int: aB < 5.81e6)
        if (!(aB < 5.5e6 && aS < 1e9)) return false;
        const double st = std::fabs(p[2]) * 0.84932180028801904272;
        const double sx = std::fabs(p[1]) * 0.84932180028801904272;
        if (!(st >= 1e-100 && st <= 1e100 && sx >= 1e-100 && sx <= 1e100)) return false;
        if (!std::isfinite(p[3] * 1.4426950408889634 / st)) return false;
        if (!std::isfinite(p[0] * p[0] / (p[1] * p[1]))) return false;
        if (!(sx * diag <= 300.0 && std::fabs(p[0]) * 0.84932180028801904272 * diag <= 300.0)) return false;
        if (split && !(st * e.window_span <= 300.0)) return false;
    }
    return true;
}

// True when every background exponent tauX^2 r^2/2 + tauT^2 dt^2/2 that the K vectors can form
// over the data's extents is at most 700: the FP64 base-2 time-sorted pass may then run
// exp2_neg_ktab, whose argument (256 log2(e) times the exponent, formed from the scaled values)
// must stay below kern::kExp2KTabMaxY = 256 x 1011 = 258,816; 700 gives 258,525, the rest a margin far
// beyond the scaled values' rounding. (The split passes' background pass spans every window of
// the culled order and keeps exp2_neg_tab.)
static bool ktab_ok(const Extents &e, const double *params, int K) {
    for (int k = 0; k < K; ++k) {
        const double *p = params + 6 * (size_t)k;
        const double aB = 0.5 * p[1] * p[1] * e.r2max + 0.5 * p[2] * p[2] * e.span * e.span;
        if (!(aB <= 700.0)) return false;
    }
    return true;
}

// Split passes (chosen automatically for likelihood passes): from kSplitMinEvents events on, when
// the locations' diagonal is at least kSplitMinRatio times the negligible-background distance
// (counted on cluster processes with windows of 32,768: ratio 3.6 forms 0.54x the background
// pairs of the time-sorted pass, ratio 7.9 0.23x, ratio 2.6 about as many).
constexpr int kSplitMinEvents = 65536;
constexpr double kSplitMinRatio = 3.0;

// Whether the split passes are expected to pay for a dataset of n events in D dimensions and its
// K vectors: the bounding box's diagonal spans at least kSplitMinRatio times the distance at which
// every background term of every vector is negligible (tauX r / sqrt(2 skip point)), so that
// groups far apart in space are skipped.
static bool split_pays(const Extents &e, const double *params, int K, int n, int D, bool fp32) {
    if (D > 3 || n < kSplitMinEvents) return false;
    const double arg = kern::background_skip_arg(n, fp32);
    for (int k = 0; k < K; ++k) {
        const double tx = std::fabs(params[6 * (size_t)k + 1]);
        if (!(tx * std::sqrt(e.r2max) >= kSplitMinRatio * std::sqrt(2.0 * arg))) return false;
    }
    return true;
}

// The culled order of the split passes, on the host: the time-sorted events cut into windows of
// cull_window(n) events, each window ordered along a Morton curve of its locations (21 bits per
// axis of the bounding box), so that every group of kern::kGroup consecutive events spans a small
// box in space and a window in time. x: D planes of npad, t: npad, perm: culled position ->
// time-sorted position (padding: 0), gstat: the groups' boxes (2D + 2 doubles each, for every
// group of the 512-event row blocks; empty groups get an inverted box that every test skips).
// locs: n x D in the caller's order; sorted: time-sorted position -> caller's index; times: in
// the caller's order.
struct CulledHost {
    std::vector<double> x, t, gstat;
    std::vector<int> perm;
};
static CulledHost make_culled(const std::vector<double> &locs, const std::vector<int> &sorted,
                              const std::vector<double> &times, int n, int D, int npad) {
    std::vector<double> lo(D), hi(D);
    for (int d = 0; d < D; ++d) {
        lo[d] = hi[d] = locs[d];
        for (int i = 1; i < n; ++i) {
            lo[d] = std::min(lo[d], locs[(size_t)i * D + d]);
            hi[d] = std::max(hi[d], locs[(size_t)i * D + d]);
        }
    }
    std::vector<unsigned long long> code(n);
    for (int s = 0; s < n; ++s) {
        const int u = sorted[s];
        unsigned long long c = 0;
        for (int d = 0; d < D; ++d) {
            const double w = hi[d] > lo[d] ? (locs[(size_t)u * D + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
            const unsigned long long q =
                (unsigned long long)std::min(2097151.0, std::max(0.0, std::floor(w * 2097152.0)));
            for (int b = 0; b < 21; ++b) c |= ((q >> b) & 1ULL) << (b * D + d);
        }
        code[s] = c;
    }
    std::vector<int> cperm(n); // culled position -> sorted position
    std::iota(cperm.begin(), cperm.end(), 0);
    const int W = cull_window(n);
    for (int w0 = 0; w0 < n; w0 += W) {
        const int w1 = std::min(n, w0 + W);
        std::stable_sort(cperm.begin() + w0, cperm.begin() + w1, [&](int a, int b) { return code[a] < code[b]; });
    }
    CulledHost c;
    c.x.resize((size_t)D * npad);
    c.t.resize(npad);
    c.perm.resize(npad);
    for (int r = 0; r < npad; ++r) {
        const int sp = cperm[r < n ? r : n - 1];
        const int u = sorted[sp];
        for (int d = 0; d < D; ++d) c.x[(size_t)d * npad + r] = locs[(size_t)u * D + d];
        c.t[r] = times[u];
        c.perm[r] = r < n ? sp : 0;
    }
    const int groups = kern::culled_groups(n);
    const int tiles = groups * kern::kGroup / kern::kTile;
    const int supers = (tiles * kern::kTile + kern::kSuperTile - 1) / kern::kSuperTile;
    const int GS = 2 * D + 2;
    c.gstat.resize((size_t)(groups + tiles + supers) * GS);
    for (int g = 0; g < groups; ++g) {
        double *e = c.gstat.data() + (size_t)g * GS;
        for (int d = 0; d <= D; ++d) {
            e[d] = 1e300;
            e[D + 1 + d] = -1e300;
        }
        for (int r = g * kern::kGroup; r < std::min(n, (g + 1) * kern::kGroup); ++r) {
            for (int d = 0; d < D; ++d) {
                e[d] = std::min(e[d], c.x[(size_t)d * npad + r]);
                e[D + 1 + d] = std::max(e[D + 1 + d], c.x[(size_t)d * npad + r]);
            }
            e[D] = std::min(e[D], c.t[r]);
            e[2 * D + 1] = std::max(e[2 * D + 1], c.t[r]);
        }
    }
    // tile and supertile boxes: unions of their groups' boxes
    auto merge = [&](int box, int g0, int g1) {
        double *e = c.gstat.data() + (size_t)box * GS;
        for (int d = 0; d <= D; ++d) {
            e[d] = 1e300;
            e[D + 1 + d] = -1e300;
        }
        for (int g = g0; g < g1; ++g) {
            const double *f = c.gstat.data() + (size_t)g * GS;
            for (int d = 0; d <= D; ++d) {
                e[d] = std::min(e[d], f[d]);
                e[D + 1 + d] = std::max(e[D + 1 + d], f[D + 1 + d]);
            }
        }
    };
    const int gpt = kern::kTile / kern::kGroup, gps = kern::kSuperTile / kern::kGroup;
    for (int u = 0; u < tiles; ++u) merge(groups + u, u * gpt, (u + 1) * gpt);
    for (int u = 0; u < supers; ++u) merge(groups + tiles + u, u * gps, std::min(groups, (u + 1) * gps));
    return c;
}

// Drives the symmetric pass over a set of device-resident datasets: plans work items, keeps the
// plan and its device tables while the shape is unchanged, owns the partial buffers.
class SymRunner {
  public:
    struct DataRef {
        const double *x, *t;
        const int *perm;
        int n, npad;
        double t_last;
    };

    ~SymRunner() {
        for (DevBuf *b : {&d_data_, &d_data_b_, &d_items_, &d_fin_, &d_res_, &d_counter_, &d_row_, &d_col_, &d_block_,
                          &d_flag_, &d_src_, &d_qbase_, &d_oscr_, &d_xs_})
            b->release();
        h_items_.release();
    }

    // Plans (if the shape changed) and enqueues the pass, finalize and reduction on stream s.
    // params: device, sum(K) x 6; result: device, sum(K) x 7; probs_dev, inv_lam_dev (may be null):
    // dataset 0, vector 0.
    // ev0/ev1/ev2 (may be null) are recorded before the pass, after it, and after the reduction.
    // sums (may be null): the per-event moments of dataset 0, vector 0 (natural units, [6][npad]).
    // Before the pass the work queue is ordered on the device by each item's expected cost for
    // these parameters, and each vector's excitation skip point formed (kern::launch_sym_order).
    void enqueue(const std::vector<DataRef> &refs, const std::vector<int> &K, int dim, const double *params,
                 bool grad, bool fp32, int rows_per_lane, int target_items, int skip_zero, int cut_s,
                 const std::vector<char> &fast, double *probs_dev, double *inv_lam_dev, double *result,
                 hipStream_t s, hipEvent_t ev0 = nullptr,
                 hipEvent_t ev1 = nullptr, hipEvent_t ev2 = nullptr, double *sums = nullptr) {
        if (fast.size() != refs.size()) throw std::invalid_argument("one base-2 flag per dataset");
        kern::SymLaunch L = prepare(refs, K, dim, grad, fp32, rows_per_lane, target_items, skip_zero, fast);
        L.cut_s = skip_zero && cut_s ? 1 : 0;
        {
            kern::SymOrder O{};
            O.src = d_src_.as<int4>();
            O.qbase = d_qbase_.as<int>();
            O.thr = d_oscr_.as<int>();
            O.key = d_oscr_.as<int>() + 2 * (size_t)onthr_;
            O.out = d_items_.as<int4>();
            O.xs = d_xs_.as<double>();
            O.nd = (int)plan_.data.size();
            O.nthr = onthr_;
            kern::launch_sym_order(L, O, params, s);
            check(hipGetLastError(), "work-queue order kernel launch");
        }
        if (ev0) check(hipEventRecord(ev0, s), "hipEventRecord");
        kern::launch_sym(L, params, grad, fp32, d_row_.as<double>(), d_col_.as<double>(), s);
        check(hipGetLastError(), "symmetric pair kernel launch");
        if (ev1) check(hipEventRecord(ev1, s), "hipEventRecord");
        kern::launch_sym_finalize(L, params, grad, d_row_.as<double>(), d_col_.as<double>(), nullptr, nullptr, sums,
                                  d_block_.as<double>(), probs_dev, inv_lam_dev, s);
        check(hipGetLastError(), "finalize kernel launch");
        kern::launch_sym_reduce(L, grad, d_block_.as<double>(), result, s);
        check(hipGetLastError(), "reduce kernel launch");
        if (ev2) check(hipEventRecord(ev2, s), "hipEventRecord");
    }

    // The split passes over one dataset: the background pass on the culled order (culled: that
    // order's events, perm = culled position -> time-sorted position, gstat = its group boxes),
    // its per-event sums into bsum (device, K x (grad ? 3 : 1) x npad), the excitation sums per
    // event of the time-sorted order into sband (same shape), then the finalize and the
    // reduction. Arguments otherwise as enqueue; ev1 is recorded after both passes.
    // host_gstat, host_params (both or neither): the culled order's box table and the parameter
    // vectors on the host, for the background pass's queue order (order_queue_culled).
    void enqueue_split(const DataRef &sorted, const DataRef &culled, const double *gstat, int K, int dim,
                       const double *params, bool grad, bool fp32, int rows_per_lane, int target_items, int skip_zero,
                       int cut_s, bool fast, double *bsum, double *sband, double *probs_dev, double *inv_lam_dev,
                       double *result, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, hipEvent_t ev2, double *sums,
                       const double *host_gstat = nullptr, const double *host_params = nullptr) {
        kern::SymLaunch L =
            prepare({sorted}, {K}, dim, grad, fp32, rows_per_lane, target_items, skip_zero, {fast ? (char)1 : (char)0});
        if (host_gstat != nullptr && host_params != nullptr && skip_zero) {
            order_queue_culled(host_gstat, host_params, dim);
            check(hipMemcpyAsync(d_items_.ptr, h_items_.ptr, sizeof(int4) * plan_.items.size(), hipMemcpyHostToDevice,
                                 s),
                  "upload work queue order");
        }
        // the background pass's table: the sorted table with the culled order's arrays
        kern::SymData b = plan_.data[0];
        b.x = culled.x;
        b.t = culled.t;
        b.perm = culled.perm;
        b.gstat = gstat;
        if (!b_valid_ || std::memcmp(&b, &b_data_, sizeof(b)) != 0) {
            d_data_b_.ensure(sizeof(kern::SymData));
            check(hipMemcpy(d_data_b_.ptr, &b, sizeof(b), hipMemcpyHostToDevice), "upload culled dataset table");
            b_data_ = b;
            b_valid_ = true;
        }
        kern::SymLaunch Lb = L;
        Lb.data = d_data_b_.as<kern::SymData>();
        Lb.pm = kern::kPassCulledB;
        if (ev0) check(hipEventRecord(ev0, s), "hipEventRecord");
        kern::launch_sym(Lb, params, grad, fp32, d_row_.as<double>(), d_col_.as<double>(), s);
        check(hipGetLastError(), "background pass launch");
        kern::launch_sym_bsums(Lb, params, grad, d_row_.as<double>(), d_col_.as<double>(), bsum, s);
        check(hipGetLastError(), "background sums launch");
        // the excitation terms of the time-sorted events, per event (FP64 in both precision modes)
        kern::Events ev{sorted.x, sorted.t, sorted.perm, sorted.n, sorted.npad, dim, sorted.t_last};
        // cut_s 0 (a probability request, or the rates of the location gradient): exact zeros only,
        // with the per-event relative stop where no moments are formed
        kern::launch_excitation_band(ev, params, K, grad, !skip_zero ? 2 : (cut_s ? 0 : (grad ? 1 : 3)), sband, s);
        check(hipGetLastError(), "excitation pass launch");
        if (ev1) check(hipEventRecord(ev1, s), "hipEventRecord");
        kern::launch_sym_finalize(L, params, grad, d_row_.as<double>(), d_col_.as<double>(), bsum, sband, sums,
                                  d_block_.as<double>(), probs_dev, inv_lam_dev, s);
        check(hipGetLastError(), "finalize kernel launch");
        kern::launch_sym_reduce(L, grad, d_block_.as<double>(), result, s);
        check(hipGetLastError(), "reduce kernel launch");
        if (ev2) check(hipEventRecord(ev2, s), "hipEventRecord");
    }

    // Finalize and reduction of one vector of one dataset from per-event moments in natural units:
    // background moments from bsum ([3 or 1][npad]), excitation moments from scache ([3 or
    // 1][npad]); the moments used are written to sums_out ([6][npad]). Arguments otherwise as
    // enqueue; ev2 is recorded after the reduction.
    // plan_grad (-1: grad): the gradient flag of the plan to use (a finalize from sums reads no
    // partial sums, so it runs on a plan made for either; reusing the current one saves re-planning).
    void enqueue_from_sums(const DataRef &ref, int dim, const double *params, bool grad, bool fp32, int rows_per_lane,
                           int target_items, const double *bsum, const double *scache, double *sums_out,
                           double *probs_dev, double *inv_lam_dev, double *result, hipStream_t s, hipEvent_t ev2,
                           int plan_grad = -1) {
        const kern::SymLaunch L =
            prepare({ref}, {1}, dim, plan_grad < 0 ? grad : plan_grad != 0, fp32, rows_per_lane, target_items, 1, {0});
        kern::launch_sym_finalize(L, params, grad, d_row_.as<double>(), d_col_.as<double>(), bsum, scache, sums_out,
                                  d_block_.as<double>(), probs_dev, inv_lam_dev, s);
        check(hipGetLastError(), "finalize kernel launch");
        kern::launch_sym_reduce(L, grad, d_block_.as<double>(), result, s);
        check(hipGetLastError(), "reduce kernel launch");
        if (ev2) check(hipEventRecord(ev2, s), "hipEventRecord");
    }

    // Finalize and reduction of K vectors of one dataset that share tauX and tauT: one set of
    // background moments for all of them (bsum, [3 or 1][npad], natural units) and per-vector
    // excitation moments (scache, [K][3 or 1][npad]); result [K][7]. No partial sums are
    // allocated. ev1 / ev2 (may be null) are recorded before the finalize and after the reduction.
    void enqueue_shared_background(const DataRef &ref, int dim, const double *params, bool grad, bool fp32, int K,
                                   const double *bsum, const double *scache, double *result, hipStream_t s,
                                   hipEvent_t ev1, hipEvent_t ev2) {
        kern::SymLaunch L = prepare({ref}, {K}, dim, grad, fp32, 2, 1, 1, {0}, false);
        L.bsum_shared = 1;
        if (ev1) check(hipEventRecord(ev1, s), "hipEventRecord");
        kern::launch_sym_finalize(L, params, grad, nullptr, nullptr, bsum, scache, nullptr, d_block_.as<double>(),
                                  nullptr, nullptr, s);
        check(hipGetLastError(), "finalize kernel launch");
        kern::launch_sym_reduce(L, grad, d_block_.as<double>(), result, s);
        check(hipGetLastError(), "reduce kernel launch");
        if (ev2) check(hipEventRecord(ev2, s), "hipEventRecord");
    }

    // Location-gradient pass over one dataset and one parameter vector (the plan of a K = 1
    // gradient evaluation, shared with it): w = 1/lambda (device, sorted), out = N x D (device).
    void enqueue_locgrad(const DataRef &ref, int dim, const double *params, const double *w, double *out,
                         int rows_per_lane, int target_items, int skip_zero, hipStream_t s) {
        const kern::SymLaunch L = prepare({ref}, {1}, dim, true, false, rows_per_lane, target_items, skip_zero, {0});
        kern::launch_sym_locgrad(L, params, w, d_row_.as<double>(), d_col_.as<double>(), out, s);
        check(hipGetLastError(), "symmetric location gradient launch");
    }

    const kern::SymPlan &plan() const { return plan_; }

    // The work queue's order for each call's parameters is formed on the device
    // (kern::launch_sym_order, enqueue with order). Items are independent (each writes its own
    // partial-sum slots, which the finalize adds in a fixed order), so the order changes the pass's
    // time and not its results. Plan order alone (by nominal size: chunk rows x columns) puts
    // blocks whose earlier columns are all within the background's reach -- every tile computed --
    // among blocks whose chunks are mostly skipped; the queue's tail then waits on the former.
    // Cost of a chunk item, in tenths of a (block rows x kTile)-pair tile: the tiles that pass the
    // workgroup-wide skip test of sym_item (the same test, on the block's first time), weighted
    // 210 where the block's first row also forms excitation terms and 100 where it forms the
    // background only; skipped tiles count 1 (a flag write). Diagonal items: rows^2 pairs of both
    // terms, 210 rows^2 / (R kTile). Both tests are monotone in the tile index (times are sorted),
    // so each (dataset, vector, block) takes two binary searches. plan_order uploads, once per
    // plan, the plan's items and each dataset's first (vector, block) index.
    void plan_order() {
        const size_t m = plan_.items.size(), nd = plan_.data.size();
        std::vector<int> qbase(nd + 1);
        size_t nthr = 0;
        for (size_t d = 0; d < nd; ++d) {
            qbase[d] = (int)nthr;
            nthr += (size_t)plan_.data[d].K * plan_.data[d].nb;
        }
        qbase[nd] = (int)nthr;
        d_src_.ensure(sizeof(int4) * m);
        check(hipMemcpy(d_src_.ptr, plan_.items.data(), sizeof(int4) * m, hipMemcpyHostToDevice), "upload plan items");
        d_qbase_.ensure(sizeof(int) * (nd + 1));
        check(hipMemcpy(d_qbase_.ptr, qbase.data(), sizeof(int) * (nd + 1), hipMemcpyHostToDevice),
              "upload dataset indices");
        d_oscr_.ensure(sizeof(int) * (2 * nthr + m));
        onthr_ = (int)nthr;
        size_t nv = 0;
        for (size_t d = 0; d < nd; ++d) nv += (size_t)plan_.data[d].K;
        d_xs_.ensure(sizeof(double) * std::max<size_t>(nv, 1));
    }

    // Forgets the culled pass's order (its box table changed).
    void reset_culled_order() { culled_order_.clear(); }

    // The same for the split passes' background pass over the culled order of dataset 0 (one
    // dataset, K vectors): there the kernel skips a (block, supertile of kSuperTile columns)
    // pair, else a (block, tile) pair, when the two bounding boxes put every background exponent
    // tauX^2 r^2 / 2 + tauT^2 dt^2 / 2 at or past the skip point (box_background_exponent). The
    // cost of a chunk item counts its tiles that pass these tests (10 each; skipped tiles 0.1,
    // skipped supertiles 0.2), with the same box tests on the host copy of the box table. The tests
    // take about 1e6 box pairs per vector at N = 1e6, so the order is kept while every vector's
    // tauX and tauT stay within 2% of the ones it was computed for (sampler steps change them
    // by far less), and recomputed otherwise.
    void order_queue_culled(const double *gstat, const double *hparams, int dim) {
        const size_t m = plan_.items.size();
        const kern::SymData &sd = plan_.data[0];
        // the cached order is a permutation of the items it was computed for: reusable while the
        // plan lists the same items (a gradient and a probability request plan the same items)
        bool reuse = culled_order_.size() == m && culled_tau_.size() == 2 * (size_t)sd.K &&
                     culled_items_.size() == m &&
                     std::memcmp(culled_items_.data(), plan_.items.data(), sizeof(int4) * m) == 0;
        for (int k = 0; reuse && k < sd.K; ++k) {
            const double *pk = hparams + 6 * (size_t)(sd.param0 + k);
            reuse = std::fabs(pk[1] / culled_tau_[2 * k] - 1.0) <= 0.02 &&
                    std::fabs(pk[2] / culled_tau_[2 * k + 1] - 1.0) <= 0.02;
        }
        if (!reuse) {
            const int R = plan_.R, T = kern::kTile, S = kern::kSuperTile, G = kern::kGroup;
            const int GS = 2 * dim + 2;
            const int box_tile0 = kern::culled_groups(sd.n);
            const int box_super0 = box_tile0 + kern::culled_groups(sd.n) * G / T;
            // merged box of `count` consecutive table entries from g
            auto box = [&](int g, int count, double *lo, double *hi) {
                for (int d = 0; d <= dim; ++d) {
                    lo[d] = gstat[(size_t)g * GS + d];
                    hi[d] = gstat[(size_t)g * GS + dim + 1 + d];
                }
                for (int q = 1; q < count; ++q)
                    for (int d = 0; d <= dim; ++d) {
                        lo[d] = std::fmin(lo[d], gstat[(size_t)(g + q) * GS + d]);
                        hi[d] = std::fmax(hi[d], gstat[(size_t)(g + q) * GS + dim + 1 + d]);
                    }
            };
            auto far = [&](const double *alo, const double *ahi, const double *blo, const double *bhi, double hx,
                           double ht) {
                double r2 = 0.0;
                for (int d = 0; d < dim; ++d) {
                    const double g = std::fmax(0.0, std::fmax(blo[d] - ahi[d], alo[d] - bhi[d]));
                    r2 = std::fma(g, g, r2);
                }
                const double gt = std::fmax(0.0, std::fmax(blo[dim] - ahi[dim], alo[dim] - bhi[dim]));
                return hx * r2 + ht * gt * gt >= sd.zero_arg_b;
            };
            cost_.assign(m, 0.0);
            double blo[kern::kMaxDim + 1], bhi[kern::kMaxDim + 1], tlo[kern::kMaxDim + 1], thi[kern::kMaxDim + 1];
            for (size_t i = 0; i < m; ++i) {
                const int4 it = plan_.items[i];
                const int P = it.y, P0 = P * R;
                if (it.z < 0) {
                    const double rows = (double)std::min(R, sd.n - P0);
                    cost_[i] = 10.0 * rows * rows / ((double)R * T);
                    continue;
                }
                const double *pk = hparams + 6 * (size_t)(sd.param0 + it.w);
                const double hx = 0.5 * pk[1] * pk[1], ht = 0.5 * pk[2] * pk[2];
                box(P0 / G, R / G, blo, bhi);
                const int c0 = it.z * sd.L, c1 = std::min(c0 + sd.L, P0);
                double c = 0.0;
                for (int j0 = c0; j0 < c1; j0 += T) {
                    if (j0 % S == 0 && j0 + S <= c1) {
                        box(box_super0 + j0 / S, 1, tlo, thi);
                        if (far(blo, bhi, tlo, thi, hx, ht)) {
                            c += 0.2;
                            j0 += S - T;
                            continue;
                        }
                    }
                    box(box_tile0 + j0 / T, 1, tlo, thi);
                    c += far(blo, bhi, tlo, thi, hx, ht) ? 0.1 : 10.0;
                }
                cost_[i] = c;
            }
            double cmax = 0.0;
            for (double c : cost_) cmax = std::max(cmax, c);
            int nbuck = 1024;
            while (nbuck < 65536 && (size_t)nbuck < m) nbuck *= 2;
            count_.assign(nbuck + 1, 0);
            key_.resize(m);
            for (size_t i = 0; i < m; ++i) {
                const int key = cmax > 0.0 ? (int)(cost_[i] / cmax * (nbuck - 1) + 0.5) : 0;
                key_[i] = nbuck - 1 - key;
                ++count_[key_[i] + 1];
            }
            for (int b = 0; b < nbuck; ++b) count_[b + 1] += count_[b];
            culled_order_.resize(m);
            for (size_t i = 0; i < m; ++i) culled_order_[count_[key_[i]]++] = plan_.items[i];
            culled_items_ = plan_.items;
            culled_tau_.resize(2 * (size_t)sd.K);
            for (int k = 0; k < sd.K; ++k) {
                culled_tau_[2 * k] = hparams[6 * (size_t)(sd.param0 + k) + 1];
                culled_tau_[2 * k + 1] = hparams[6 * (size_t)(sd.param0 + k) + 2];
            }
        }
        h_items_.ensure(sizeof(int4) * m);
        std::memcpy(h_items_.ptr, culled_order_.data(), sizeof(int4) * m);
    }

  private:
    // Plans (if the shape changed), uploads the tables (if they changed) and fills the launch.
    // fast[d]: dataset d runs the base-2 form (1), or the base-2 form with every background
    // exponent in exp2_neg_ktab's range (2, ktab_ok; used in FP64 when every dataset has it).
    kern::SymLaunch prepare(const std::vector<DataRef> &refs, const std::vector<int> &K, int dim, bool grad,
                            bool fp32, int rows_per_lane, int target_items, int skip_zero,
                            const std::vector<char> &fast, bool partials = true) {
        std::vector<int> n(refs.size());
        for (size_t d = 0; d < refs.size(); ++d) n[d] = refs[d].n;
        const bool same = valid_ && n == n_ && K == K_ && rows_per_lane == A_ && target_items == target_ && grad == grad_;
        if (!same) {
            plan_ = kern::plan_sym(n, K, rows_per_lane, target_items, grad);
            n_ = n;
            K_ = K;
            A_ = rows_per_lane;
            target_ = target_items;
            grad_ = grad;
            d_items_.ensure(sizeof(int4) * plan_.items.size());
            check(hipMemcpy(d_items_.ptr, plan_.items.data(), sizeof(int4) * plan_.items.size(), hipMemcpyHostToDevice),
                  "upload work items");
            plan_order();
            d_fin_.ensure(sizeof(int4) * plan_.fin_items.size());
            check(hipMemcpy(d_fin_.ptr, plan_.fin_items.data(), sizeof(int4) * plan_.fin_items.size(),
                            hipMemcpyHostToDevice),
                  "upload finalize items");
            d_res_.ensure(sizeof(int2) * plan_.res_items.size());
            check(hipMemcpy(d_res_.ptr, plan_.res_items.data(), sizeof(int2) * plan_.res_items.size(),
                            hipMemcpyHostToDevice),
                  "upload result items");
            if (!d_counter_.ptr) {
                d_counter_.ensure(sizeof(int));
                check(hipMemset(d_counter_.ptr, 0, sizeof(int)), "zero queue counter");
            }
            d_block_.ensure(sizeof(double) * std::max<size_t>(plan_.block_partials, 1) * kern::kResultWidth * 2);
            refs_.clear();
            valid_ = true;
            b_valid_ = false;
        }
        // the pass's partial sums (a finalize from per-event sums reads none)
        if (partials) {
            d_row_.ensure(sizeof(double) * std::max<size_t>(plan_.rowpart_doubles, 1));
            d_col_.ensure(sizeof(double) * std::max<size_t>(plan_.colpart_doubles, 1));
            d_flag_.ensure(std::max<size_t>(plan_.colflag_bytes, 1));
        }
        bool refs_same = refs_.size() == refs.size();
        for (size_t d = 0; refs_same && d < refs.size(); ++d)
            refs_same = refs_[d].x == refs[d].x && refs_[d].t == refs[d].t && refs_[d].perm == refs[d].perm &&
                        refs_[d].npad == refs[d].npad && refs_[d].t_last == refs[d].t_last;
        for (size_t d = 0; refs_same && d < refs.size(); ++d)
            refs_same = plan_.data[d].zero_arg_b == kern::background_skip_arg(refs[d].n, fp32) &&
                        plan_.data[d].fast == (fast[d] ? 1 : 0);
        if (!refs_same) {
            for (size_t d = 0; d < refs.size(); ++d) {
                kern::SymData &sd = plan_.data[d];
                sd.zero_arg_b = kern::background_skip_arg(refs[d].n, fp32);
                sd.fast = fast[d] ? 1 : 0;
                sd.x = refs[d].x;
                sd.t = refs[d].t;
                sd.perm = refs[d].perm;
                sd.npad = refs[d].npad;
                sd.t_last = refs[d].t_last;
            }
            d_data_.ensure(sizeof(kern::SymData) * plan_.data.size());
            check(hipMemcpy(d_data_.ptr, plan_.data.data(), sizeof(kern::SymData) * plan_.data.size(),
                            hipMemcpyHostToDevice),
                  "upload dataset table");
            refs_ = refs;
        }
        if (cus_ == 0) {
            int dev = 0;
            check(hipGetDevice(&dev), "hipGetDevice");
            check(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev), "device CU count");
        }
        kern::SymLaunch L{};
        L.data = d_data_.as<kern::SymData>();
        L.item = d_items_.as<int4>();
        L.items = (int)plan_.items.size();
        L.counter = d_counter_.as<int>();
        L.xs = d_xs_.as<double>(); // written by the order kernel before a pass in time order
        L.grid = std::min(L.items, std::max(cus_, 1) * 4);
        L.dim = dim;
        L.A = plan_.A;
        L.R = plan_.R;
        L.skip_zero = skip_zero;
        L.zero_arg_s = kern::kZeroExpArg;
        L.cut_s = 0;
        int nfast = 0;
        for (char f : fast) nfast += f ? 1 : 0;
        L.fast = nfast == 0 ? 0 : (nfast == (int)fast.size() ? 1 : 2);
        L.b2tab = fp32 ? 0 : 1;
        int nktab = 0;
        for (char f : fast) nktab += f == 2 ? 1 : 0;
        L.ktab = !fp32 && L.fast == 1 && nktab == (int)fast.size() ? 1 : 0;
        L.fin_item = d_fin_.as<int4>();
        L.fin_items = (int)plan_.fin_items.size();
        L.res_item = d_res_.as<int2>();
        L.res_items = (int)plan_.res_items.size();
        L.colflag = d_flag_.as<unsigned char>();
        return L;
    }

    kern::SymPlan plan_;
    bool valid_ = false, grad_ = false;
    std::vector<int> n_, K_;
    int A_ = 0, target_ = 0, cus_ = 0;
    std::vector<DataRef> refs_;
    kern::SymData b_data_{};
    bool b_valid_ = false;
    DevBuf d_data_, d_data_b_, d_items_, d_fin_, d_res_, d_counter_, d_row_, d_col_, d_block_, d_flag_;
    // the device work-queue order (plan_order, enqueue with order): the plan's items, the
    // datasets' first (vector, block) indices and scratch (thresholds, then the items' classes)
    DevBuf d_src_, d_qbase_, d_oscr_, d_xs_;
    int onthr_ = 0;
    // the culled pass's order (order_queue_culled): pinned staging buffer and scratch
    HostBuf h_items_;
    std::vector<int> count_, key_;
    std::vector<double> cost_;
    // order_queue_culled: the last order, the plan items and the (tauX, tauT) per vector it was
    // computed for
    std::vector<int4> culled_order_, culled_items_;
    std::vector<double> culled_tau_;
};

// Pair-pass configuration chosen by the autotuner.
struct TuneConfig {
    int kind = HPH_PAIR_KERNEL_ROWWISE; // HPH_PAIR_KERNEL_ROWWISE or HPH_PAIR_KERNEL_SYMMETRIC
    int rows_per_lane = 2;              // symmetric pass: event rows per lane
    int target_blocks = 4096;           // grid size the work decomposition aims for
};

// Shape of a tuned problem: log2 of N, dimension, batch size, gradient, precision mode; for a
// dataset batch (datasets > 0) log2n is log2 of the launch's pair-evaluations sum_d K_d N_d^2 and K
// the total number of vectors.
struct TuneKey {
    int log2n = 0, dim = 0, K = 0, grad = 0, fp32 = 0, datasets = 0;
    bool operator<(const TuneKey &o) const {
        return std::tie(log2n, dim, K, grad, fp32, datasets) < std::tie(o.log2n, o.dim, o.K, o.grad, o.fp32, o.datasets);
    }
};

// Process-wide tuning results, shared by all engines and dataset batches. When HPH_TUNE_CACHE names
// a file, results are read from it at first use and appended to it after each tuning run (one line
// per shape: "log2n dim K grad fp32 datasets kind rows_per_lane target_blocks").
class TuneCache {
  public:
    bool lookup(const TuneKey &k, TuneConfig *out) {
        std::lock_guard<std::mutex> lock(mu_);
        load();
        auto it = map_.find(k);
        if (it == map_.end()) return false;
        *out = it->second;
        return true;
    }
    void store(const TuneKey &k, const TuneConfig &c) {
        std::lock_guard<std::mutex> lock(mu_);
        load();
        map_[k] = c;
        const char *path = std::getenv("HPH_TUNE_CACHE");
        if (path && *path) {
            if (FILE *f = std::fopen(path, "a")) {
                std::fprintf(f, "%d %d %d %d %d %d %d %d %d\n", k.log2n, k.dim, k.K, k.grad, k.fp32, k.datasets,
                             c.kind, c.rows_per_lane, c.target_blocks);
                std::fclose(f);
            }
        }
    }

  private:
    void load() {
        if (loaded_) return;
        loaded_ = true;
        const char *path = std::getenv("HPH_TUNE_CACHE");
        if (!path || !*path) return;
        FILE *f = std::fopen(path, "r");
        if (!f) return;
        TuneKey k;
        TuneConfig c;
        while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d", &k.log2n, &k.dim, &k.K, &k.grad, &k.fp32, &k.datasets,
                           &c.kind, &c.rows_per_lane, &c.target_blocks) == 9) {
            const bool ok = (c.kind == HPH_PAIR_KERNEL_ROWWISE || c.kind == HPH_PAIR_KERNEL_SYMMETRIC) &&
                            (c.rows_per_lane == 1 || c.rows_per_lane == 2) && c.target_blocks >= 64 &&
                            c.target_blocks <= (1 << 24) && (c.kind == HPH_PAIR_KERNEL_ROWWISE || k.dim <= 3) &&
                            k.datasets >= 0;
            if (ok) map_[k] = c;
        }
        std::fclose(f);
    }
    std::mutex mu_;
    bool loaded_ = false;
    std::map<TuneKey, TuneConfig> map_;
};

static TuneCache &tune_cache() {
    static TuneCache cache;
    return cache;
}

class HipHawkes : public HawkesEngine {
  public:
    HipHawkes(int D, int N, long flags, int device)
        : HawkesEngine(D, N, flags), device_(device < 0 ? 0 : device), fp32_((flags & FLOAT) != 0) {
        if (D < 1 || D > kern::kMaxDim) throw std::invalid_argument("embeddingDimension must be in 1..8");
        if (N < 1) throw std::invalid_argument("locationCount must be >= 1");
        if (N > (1 << 30)) throw std::invalid_argument("locationCount must be at most 2^30");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device_ >= count) throw std::invalid_argument("device index out of range");
        check(hipSetDevice(device_), "hipSetDevice");
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        npad_ = ((N + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock) * kern::kRowsPerBlock;
        locs_.assign((size_t)N * D, 0.0);
        times_.assign((size_t)N, 0.0);
        d_x_.ensure(sizeof(double) * (size_t)D * npad_);
        d_t_.ensure(sizeof(double) * (size_t)npad_);
        d_perm_.ensure(sizeof(int) * (size_t)npad_);
        d_probs_.ensure(sizeof(double) * (size_t)N);
    }

    ~HipHawkes() override {
        (void)hipSetDevice(device_);
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (DevBuf *b : {&d_x_, &d_t_, &d_perm_, &d_probs_, &d_params_, &d_part_, &d_block_, &d_result_, &d_w_,
                          &d_locgrad_, &d_xc_, &d_tc_, &d_cperm_, &d_gstat_, &d_bsum_, &d_sums_[0], &d_sums_[1],
                          &d_sband_, &d_sband_batch_})
            b->release();
        h_stage_.release();
        h_result_.release();
        h_params_.release();
        for (auto &e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    int getInternalDimension() override { return embeddingDimension; }

    void updateLocations(int index, double *location, size_t length) override {
        const int D = embeddingDimension;
        if (location == nullptr) throw std::invalid_argument("locations pointer is null");
        if (index == -1) {
            if (length != (size_t)D * locationCount)
                throw std::invalid_argument("updateLocations(-1): length must be locationCount * dimension");
            for (size_t q = 0; q < length; ++q)
                if (!std::isfinite(location[q])) throw std::invalid_argument("locations must be finite");
            std::memcpy(locs_.data(), location, sizeof(double) * length);
            have_locs_ = true;
        } else {
            if (index < 0 || index >= locationCount) throw std::invalid_argument("location index out of range");
            if (length != (size_t)D) throw std::invalid_argument("updateLocations(i): length must be dimension");
            for (size_t q = 0; q < length; ++q)
                if (!std::isfinite(location[q])) throw std::invalid_argument("locations must be finite");
            std::memcpy(locs_.data() + (size_t)index * D, location, sizeof(double) * D);
            updatedLocation = index;
        }
        data_dirty_ = true;
        invalidate();
    }

    void setTimesData(double *data, size_t length) override {
        if (data == nullptr) throw std::invalid_argument("times pointer is null");
        if (length != (size_t)locationCount) throw std::invalid_argument("setTimesData: length must be locationCount");
        for (size_t q = 0; q < length; ++q)
            if (!std::isfinite(data[q])) throw std::invalid_argument("times must be finite");
        std::memcpy(times_.data(), data, sizeof(double) * length);
        have_times_ = true;
        data_dirty_ = true;
        order_dirty_ = true;
        invalidate();
    }

    void setParameters(double *data, size_t length) override {
        if (data == nullptr) throw std::invalid_argument("parameters pointer is null");
        if (length != HPH_PARAMETER_COUNT) throw std::invalid_argument("setParameters: length must be 6");
        for (int q = 0; q < 6; ++q)
            if (!std::isfinite(data[q])) throw std::invalid_argument("parameters must be finite");
        std::memcpy(params_, data, sizeof(params_));
        have_params_ = true;
        invalidate();
    }

    double getSumOfLikContribs() override {
        if (!ll_valid_) evaluate_single(false, nullptr);
        return ll_;
    }

    void getLogLikeGradient(double *out, size_t length) override {
        if (out == nullptr) throw std::invalid_argument("gradient pointer is null");
        if (length != HPH_PARAMETER_COUNT) throw std::invalid_argument("gradient length must be 6");
        if (!grad_valid_) evaluate_single(true, nullptr);
        std::memcpy(out, grad_, sizeof(grad_));
    }

    double getLogLikelihoodAndGradient(double *gradient6) override {
        if (gradient6 == nullptr) throw std::invalid_argument("gradient pointer is null");
        if (!grad_valid_) evaluate_single(true, nullptr);
        std::memcpy(gradient6, grad_, sizeof(grad_));
        return ll_;
    }

    // A pass that skips exact excitation zeros only (evaluate_single with a probability output),
    // unless the last evaluation of the unchanged state already was one (its probabilities are
    // kept); the reference always recomputes (src/NewHawkes.hpp:279 notes the missing caching).
    void getProbsSelfExcite(double *result, size_t length) override {
        if (result == nullptr) throw std::invalid_argument("result pointer is null");
        if (length != (size_t)locationCount) throw std::invalid_argument("getProbsSelfExcite: length must be locationCount");
        if (probs_valid_) {
            check(hipSetDevice(device_), "hipSetDevice");
            check(hipMemcpy(result, d_probs_.ptr, sizeof(double) * locationCount, hipMemcpyDeviceToHost),
                  "download probs");
            return;
        }
        if (probs_from_cache(result)) return;
        evaluate_single(false, result);
    }

    // w = 1/lambda from the last single-vector evaluation of the current state (else a natural-form
    // log-likelihood pass is run, which also caches the log-likelihood; FP32-bulk engines run a
    // row-wise FP64 rate pass instead), then the location-gradient pass of the engine's pass
    // configuration (symmetric for D <= 3, else row-wise); the result is downloaded in the
    // caller's event order. Timing: [0] = the rate pass run by this call (0
    // when reused), [1] = the location-gradient pass.
    void getLocationGradient(double *out, size_t length) override {
        if (out == nullptr) throw std::invalid_argument("location gradient pointer is null");
        const int N = locationCount, D = embeddingDimension;
        if (length != (size_t)N * D) throw std::invalid_argument("location gradient length must be N*D");
        require_ready();
        if (!have_params_) throw NotInitialized("parameters are not set (setParameters)");
        check(hipSetDevice(device_), "hipSetDevice");
        double rate_ms = 0.0;
        const int skip_zero = (flags & ALL_PAIRS) ? 0 : 1;
        if (!w_valid_ && !fp32_) {
            force_natural_ = true;
            try {
                evaluate_single(false, nullptr);
            } catch (...) {
                force_natural_ = false;
                throw;
            }
            force_natural_ = false;
            rate_ms = last_ms_[0] + last_ms_[1];
        }
        if (data_dirty_) upload();
        d_params_.ensure(sizeof(double) * 6);
        h_params_.ensure(sizeof(double) * 6);
        std::memcpy(h_params_.ptr, params_, sizeof(params_));
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6, hipMemcpyHostToDevice, stream_),
              "upload params");
        const kern::Events ev = events();
        kern::PairGrid g = kern::plan_pair_grid(N, 1, 4096);
        g.skip_zero = skip_zero;
        d_part_.ensure(sizeof(double) * (size_t)g.chunks * std::max(2, D) * npad_);
        d_w_.ensure(sizeof(double) * npad_);
        d_locgrad_.ensure(sizeof(double) * (size_t)N * D);
        if (fp32_) {
            if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            kern::launch_pair(ev, d_params_.as<double>(), 1, g, false, false, d_part_.as<double>(), stream_);
            check(hipGetLastError(), "rate pass launch");
            kern::launch_inv_rate(ev, d_params_.as<double>(), g, d_part_.as<double>(), d_w_.as<double>(), stream_);
            check(hipGetLastError(), "inverse rate launch");
        } else if (timing_) {
            check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
        }
        if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
        const TuneConfig cfg = config_for(1, true);
        if ((cfg.kind == HPH_PAIR_KERNEL_SYMMETRIC || cfg.kind == HPH_PAIR_KERNEL_SPLIT) && D <= 3 &&
            sym_fits(1, cfg, true)) {
            SymRunner::DataRef ref{ev.x, ev.t, ev.perm, ev.n, ev.npad, ev.t_last};
            sym_.enqueue_locgrad(ref, D, d_params_.as<double>(), d_w_.as<double>(), d_locgrad_.as<double>(),
                                 cfg.rows_per_lane, cfg.target_blocks, skip_zero, stream_);
        } else {
            kern::launch_locgrad(ev, d_params_.as<double>(), d_w_.as<double>(), g, d_part_.as<double>(),
                                 d_locgrad_.as<double>(), stream_);
            check(hipGetLastError(), "location gradient launch");
        }
        if (timing_) check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
        check(hipMemcpyAsync(out, d_locgrad_.ptr, sizeof(double) * (size_t)N * D, hipMemcpyDeviceToHost, stream_),
              "download location gradient");
        check(hipStreamSynchronize(stream_), "location gradient sync");
        if (timing_) {
            record_timing();
            if (!fp32_) last_ms_[0] = rate_ms;
        }
    }

    // Snapshot / restore of parameters, locations and the last log-likelihood: the intent of
    // src/NewHawkes.hpp:284-318 (whose storeState copies in the wrong direction).
    void storeState() override {
        std::memcpy(stored_params_, params_, sizeof(params_));
        stored_have_params_ = have_params_;
        stored_locs_ = locs_;
        stored_have_locs_ = have_locs_;
        stored_ll_ = ll_;
        stored_ll_valid_ = ll_valid_;
        has_stored_ = true;
    }

    void restoreState() override {
        if (!has_stored_) return;
        std::memcpy(params_, stored_params_, sizeof(params_));
        have_params_ = stored_have_params_;
        if (stored_locs_ != locs_) {
            locs_ = stored_locs_;
            data_dirty_ = true;
        }
        have_locs_ = stored_have_locs_;
        invalidate();
        ll_ = stored_ll_;
        ll_valid_ = stored_ll_valid_;
        updatedLocation = -1;
    }

    // The reference uses acceptState to fold a single-location update into cached pairwise
    // data (src/NewHawkes.hpp:293-303); this engine keeps no pairwise cache.
    void acceptState() override { updatedLocation = -1; }

    void setParametersBatch(const double *params, size_t count) override {
        if (params == nullptr) throw std::invalid_argument("parameters pointer is null");
        if (count == 0 || count % HPH_PARAMETER_COUNT != 0)
            throw std::invalid_argument("batch parameter count must be a positive multiple of 6");
        if (count / HPH_PARAMETER_COUNT > 65535) throw std::invalid_argument("at most 65535 parameter vectors");
        for (size_t q = 0; q < count; ++q)
            if (!std::isfinite(params[q])) throw std::invalid_argument("parameters must be finite");
        batch_.assign(params, params + count);
    }

    void getLogLikelihoodBatch(double *loglik, double *gradient, size_t K) override {
        if (loglik == nullptr) throw std::invalid_argument("loglik pointer is null");
        if (batch_.empty()) throw NotInitialized("batch parameters are not set");
        if (K != batch_.size() / HPH_PARAMETER_COUNT)
            throw std::invalid_argument("K does not match the number of batch parameter vectors");
        const bool grad = gradient != nullptr;
        if (evaluate_batch_shared((int)K, grad)) {
            last_eval_ = HPH_EVAL_EXCITATION;
        } else {
            run(batch_.data(), (int)K, grad, nullptr);
            last_eval_ = HPH_EVAL_FULL;
        }
        const double *r = h_result_.as<double>();
        for (size_t k = 0; k < K; ++k) {
            loglik[k] = r[k * kern::kResultWidth];
            if (grad)
                for (int q = 0; q < 6; ++q) gradient[k * 6 + q] = r[k * kern::kResultWidth + 1 + q];
        }
    }

    void setTiming(bool enabled) override {
        check(hipSetDevice(device_), "hipSetDevice");
        // the events only time the stream's work (read after a stream synchronisation): no
        // system-scope cache write-back at each of them
        if (enabled && !ev_[0])
            for (auto &e : ev_) check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
        timing_ = enabled;
    }

    // Times every candidate configuration on the current data and parameters (LL + gradient for
    // K = batch size, or 1) and keeps the fastest for this shape in the process-wide cache.
    void autotune(int repeats, double *best_ms) override {
        if (!have_params_ && batch_.empty()) throw NotInitialized("parameters are not set (setParameters)");
        if (repeats < 1) repeats = 1;
        const bool use_batch = !batch_.empty();
        const int K = use_batch ? (int)(batch_.size() / 6) : 1;
        const double *params = use_batch ? batch_.data() : params_;
        std::vector<TuneConfig> cands;
        for (int target : {1024, 2048, 4096, 8192, 65536}) {
            TuneConfig c;
            c.kind = HPH_PAIR_KERNEL_ROWWISE;
            c.target_blocks = target;
            cands.push_back(c);
            if (target > 8192) continue; // only the row-wise pass changes layout (narrow) at 65536
            if (embeddingDimension <= 3 && kernel_ != HPH_PAIR_KERNEL_ROWWISE) {
                for (int A : {1, 2}) {
                    TuneConfig d;
                    d.kind = HPH_PAIR_KERNEL_SYMMETRIC;
                    d.rows_per_lane = A;
                    d.target_blocks = target;
                    cands.push_back(d);
                }
            }
        }
        if (kernel_ == HPH_PAIR_KERNEL_SYMMETRIC || kernel_ == HPH_PAIR_KERNEL_SPLIT) {
            std::vector<TuneConfig> only;
            for (auto &c : cands)
                if (c.kind == HPH_PAIR_KERNEL_SYMMETRIC) only.push_back(c);
            cands.swap(only);
        }
        const bool was_timing = timing_;
        setTiming(true);
        TuneConfig best;
        double best_t = 1e300;
        for (const TuneConfig &c : cands) {
            run_with(params, K, true, nullptr, c); // warm: plans, buffers, code objects
            double t = 1e300;
            for (int r = 0; r < repeats; ++r) {
                run_with(params, K, true, nullptr, c);
                t = std::min(t, last_ms_[0] + last_ms_[1]);
            }
            if (t < best_t) {
                best_t = t;
                best = c;
            }
        }
        timing_ = was_timing;
        tune_cache().store(tune_key(K, true), best);
        tune_cache().store(tune_key(K, false), best);
        if (best_ms) *best_ms = best_t;
        tuned_ = best;
    }

    void getTunedConfig(int *out, size_t length) override {
        if (out == nullptr || length < 3) throw std::invalid_argument("tuned-config output needs 3 values");
        TuneConfig c = config_for(1, true);
        if (c.kind == HPH_PAIR_KERNEL_SYMMETRIC && kernel_ == HPH_PAIR_KERNEL_AUTO && !(flags & ALL_PAIRS) &&
            have_params_ && !data_dirty_ && split_pays(params_, 1))
            c.kind = HPH_PAIR_KERNEL_SPLIT;
        out[0] = c.kind;
        out[1] = c.rows_per_lane;
        out[2] = c.target_blocks;
    }

    void setPairKernel(int kind) override { setPairConfig(kind, 2, 4096); }

    void setPairConfig(int kind, int rows_per_lane, int target_blocks) override {
        if (kind < HPH_PAIR_KERNEL_AUTO || kind > HPH_PAIR_KERNEL_SPLIT)
            throw std::invalid_argument("unknown pair kernel kind");
        if ((kind == HPH_PAIR_KERNEL_SYMMETRIC || kind == HPH_PAIR_KERNEL_SPLIT) && embeddingDimension > 3)
            throw std::invalid_argument("the symmetric pair kernel is built for dimensions 1..3");
        if (rows_per_lane != 1 && rows_per_lane != 2) throw std::invalid_argument("rows per lane must be 1 or 2");
        if (target_blocks < 64 || target_blocks > (1 << 20))
            throw std::invalid_argument("target grid size must be in 64..2^20");
        kernel_ = kind;
        forced_.kind = kind == HPH_PAIR_KERNEL_AUTO ? HPH_PAIR_KERNEL_ROWWISE : kind;
        forced_.rows_per_lane = rows_per_lane;
        forced_.target_blocks = target_blocks;
    }

    int getLastEvaluation() override { return last_eval_; }

    void getLastPass(int *out, size_t length) override {
        if (out == nullptr || length < 2) throw std::invalid_argument("last-pass output needs 2 values");
        if (last_kind_ == 0) throw NotInitialized("no pair pass has run yet");
        out[0] = last_kind_;
        out[1] = last_fast_ ? 1 : 0;
        if (length >= 3) out[2] = last_kind_ == HPH_PAIR_KERNEL_ROWWISE ? last_k_ : last_sub_batch_;
        if (length >= 4) out[3] = last_ktab_ ? 1 : 0;
    }

    // The six per-event sums of the current parameters' last single-vector evaluation with the
    // gradient (the moments every later re-evaluation starts from), natural units, caller's order.
    void getEventSums(double *out, size_t length) override {
        const int N = locationCount;
        if (out == nullptr || length != (size_t)N * 6) throw std::invalid_argument("event sums need N x 6 values");
        if (!cache_valid_ || !cache_grad_ || data_dirty_ || std::memcmp(cache_params_, params_, sizeof(params_)) != 0)
            throw NotInitialized("no per-event sums of the current parameters: evaluate the log-likelihood and "
                                 "gradient first (dimensions 1..3)");
        check(hipSetDevice(device_), "hipSetDevice");
        std::vector<double> host((size_t)6 * npad_);
        check(hipMemcpy(host.data(), d_sums_[sums_cur_].ptr, sizeof(double) * host.size(), hipMemcpyDeviceToHost),
              "download event sums");
        for (int s = 0; s < N; ++s)
            for (int q = 0; q < 6; ++q) out[(size_t)perm_[s] * 6 + q] = host[(size_t)q * npad_ + s];
    }

    void getTiming(double *out, size_t length) override {
        if (out == nullptr || length < 2) throw std::invalid_argument("timing output needs 2 values");
        out[0] = last_ms_[0];
        out[1] = last_ms_[1];
    }

  private:
    void invalidate() { ll_valid_ = grad_valid_ = probs_valid_ = w_valid_ = false; }

    void require_ready() const {
        if (!have_locs_) throw NotInitialized("locations are not set (updateLocations)");
        if (!have_times_) throw NotInitialized("times are not set (setTimesData)");
    }

    // Sorts by time (stable) and uploads the structure-of-arrays copy.
    void upload() {
        const int N = locationCount, D = embeddingDimension;
        if (order_dirty_) {
            perm_.resize(N);
            std::iota(perm_.begin(), perm_.end(), 0);
            std::stable_sort(perm_.begin(), perm_.end(), [&](int a, int b) { return times_[a] < times_[b]; });
            order_dirty_ = false;
        }
        const size_t bytes = sizeof(double) * (size_t)(D + 1) * npad_ + sizeof(int) * (size_t)npad_;
        h_stage_.ensure(bytes);
        double *hx = h_stage_.as<double>();
        double *ht = hx + (size_t)D * npad_;
        int *hp = reinterpret_cast<int *>(ht + npad_);
        for (int s = 0; s < npad_; ++s) {
            const int u = perm_[s < N ? s : N - 1];
            for (int d = 0; d < D; ++d) hx[(size_t)d * npad_ + s] = locs_[(size_t)u * D + d];
            ht[s] = times_[u];
            hp[s] = u;
        }
        check(hipMemcpyAsync(d_x_.ptr, hx, sizeof(double) * (size_t)D * npad_, hipMemcpyHostToDevice, stream_),
              "upload x");
        check(hipMemcpyAsync(d_t_.ptr, ht, sizeof(double) * npad_, hipMemcpyHostToDevice, stream_), "upload t");
        check(hipMemcpyAsync(d_perm_.ptr, hp, sizeof(int) * npad_, hipMemcpyHostToDevice, stream_), "upload perm");
        check(hipStreamSynchronize(stream_), "upload sync");
        extents_ = extents_of(locs_, ht, (size_t)N, D);
        data_dirty_ = false;
        culled_valid_ = false;
        cache_valid_ = false;
    }

    // The culled order of the split passes (make_culled), uploaded with its group boxes.
    void build_culled() {
        const int N = locationCount, D = embeddingDimension;
        const CulledHost c = make_culled(locs_, perm_, times_, N, D, npad_);
        d_xc_.ensure(sizeof(double) * (size_t)D * npad_);
        d_tc_.ensure(sizeof(double) * (size_t)npad_);
        d_cperm_.ensure(sizeof(int) * (size_t)npad_);
        d_gstat_.ensure(sizeof(double) * c.gstat.size());
        check(hipMemcpy(d_xc_.ptr, c.x.data(), sizeof(double) * c.x.size(), hipMemcpyHostToDevice), "upload culled x");
        check(hipMemcpy(d_tc_.ptr, c.t.data(), sizeof(double) * c.t.size(), hipMemcpyHostToDevice), "upload culled t");
        check(hipMemcpy(d_cperm_.ptr, c.perm.data(), sizeof(int) * c.perm.size(), hipMemcpyHostToDevice),
              "upload culled order");
        check(hipMemcpy(d_gstat_.ptr, c.gstat.data(), sizeof(double) * c.gstat.size(), hipMemcpyHostToDevice),
              "upload group boxes");
        gstat_host_ = c.gstat;
        sym_.reset_culled_order();
        culled_valid_ = true;
    }

    bool split_pays(const double *params, int K) const {
        return ::hph::split_pays(extents_, params, K, locationCount, embeddingDimension, fp32_);
    }

    kern::Events events() const {
        kern::Events ev;
        ev.x = d_x_.as<double>();
        ev.t = d_t_.as<double>();
        ev.perm = d_perm_.as<int>();
        ev.n = locationCount;
        ev.npad = npad_;
        ev.dim = embeddingDimension;
        ev.t_last = times_[locationCount - 1];
        return ev;
    }

    // Pair-pass configuration for K parameter vectors: the autotuned one when this engine or an
    // earlier one in the process has tuned this shape, else the built-in choice (row-wise with the
    // narrow layout for small N or D > 3, the symmetric pass above; thresholds below).
    TuneConfig config_for(int K, bool grad) const {
        TuneConfig c;
        const bool sym_ok = embeddingDimension <= 3;
        if (kernel_ != HPH_PAIR_KERNEL_AUTO) {
            c = forced_;
        } else if (!sym_ok) {
            c.kind = HPH_PAIR_KERNEL_ROWWISE;
            if (locationCount < kNarrowMaxEvents) c.target_blocks = 65536;
        } else {
            TuneConfig t;
            if (tune_cache().lookup(tune_key(K, grad), &t)) return t;
            if (locationCount < kNarrowMaxEvents) {
                c.kind = HPH_PAIR_KERNEL_ROWWISE;
                c.target_blocks = 65536; // the narrow layout below ~32k events
            } else {
                c.kind = HPH_PAIR_KERNEL_SYMMETRIC;
                c.rows_per_lane = locationCount < kSymA1MaxEvents ? 1 : 2;
                // from the split passes' size on, finer items (N = 1e6, 100 km, LL + gradient:
                // 25.8 ms at target 8192 against 27.1 ms at 4096; FP32-bulk 18.2 against 19.8)
                if (locationCount >= kSplitMinEvents) c.target_blocks = 8192;
            }
        }
        return c;
    }

    // Whether the symmetric pass's partial sums (row partials per chunk, column partials per
    // (block, earlier column): ~3 N^2 / (2 R) doubles per vector with gradient) fit in 80% of the
    // device memory; memoised per shape.
    bool sym_fits(int K, const TuneConfig &cfg, bool grad) {
        const long long key = (((long long)K * 4 + cfg.rows_per_lane) * (1LL << 21) + cfg.target_blocks) * 2 + (grad ? 1 : 0);
        auto it = sym_fits_.find(key);
        if (it != sym_fits_.end()) return it->second;
        const kern::SymPlan plan = kern::plan_sym({locationCount}, {K}, cfg.rows_per_lane, cfg.target_blocks, grad);
        const double bytes = 8.0 * ((double)plan.rowpart_doubles + (double)plan.colpart_doubles +
                                    (double)plan.block_partials * kern::kResultWidth * 2);
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const bool fits = bytes <= 0.8 * (double)total_b;
        sym_fits_[key] = fits;
        return fits;
    }

    TuneKey tune_key(int K, bool grad) const {
        TuneKey key;
        int b = 0;
        while ((1 << (b + 1)) <= locationCount) ++b;
        key.log2n = b;
        key.dim = embeddingDimension;
        key.K = K;
        key.grad = grad ? 1 : 0;
        key.fp32 = fp32_ ? 1 : 0;
        return key;
    }

    // Runs the pair pass, finalize and reduction kernels for K parameter vectors; results land in
    // h_result_ ([K][7]).
    // probs_host: also download the probabilities (implies write_probs); write_probs: the finalize
    // kernel writes vector 0's probabilities to d_probs_.
    // sums (may be null; K = 1): receives the per-event moments ([6][npad], natural units).
    void run(const double *params, int K, bool grad, double *probs_host, bool write_probs = false,
             double *sums = nullptr) {
        run_with(params, K, grad, probs_host, config_for(K, grad), write_probs, sums);
    }

    void run_with(const double *params, int K, bool grad, double *probs_host, const TuneConfig &cfg_in,
                  bool write_probs = false, double *sums = nullptr) {
        require_ready();
        check(hipSetDevice(device_), "hipSetDevice");
        if (data_dirty_) upload();
        d_params_.ensure(sizeof(double) * 6 * (size_t)K);
        const int fin_blocks = (locationCount + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock;
        d_block_.ensure(sizeof(double) * (size_t)K * fin_blocks * kern::kResultWidth * 2);
        d_result_.ensure(sizeof(double) * (size_t)K * kern::kResultWidth);
        h_result_.ensure(sizeof(double) * (size_t)K * kern::kResultWidth);
        h_params_.ensure(sizeof(double) * 6 * (size_t)K);
        std::memcpy(h_params_.ptr, params, sizeof(double) * 6 * (size_t)K);
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6 * (size_t)K, hipMemcpyHostToDevice,
                             stream_),
              "upload params");
        const kern::Events ev = events();
        const int skip_zero = (flags & ALL_PAIRS) ? 0 : 1;
        double *probs_dev = (probs_host || write_probs) ? d_probs_.as<double>() : nullptr;
        TuneConfig cfg = cfg_in;
        const bool sym_kind = cfg.kind == HPH_PAIR_KERNEL_SYMMETRIC || cfg.kind == HPH_PAIR_KERNEL_SPLIT;
        // K vectors whose symmetric partial sums do not fit run as consecutive sub-batches that do
        // (the row-wise pass, the other way out, forms every pair of every vector without culling)
        int sub = K;
        if (sym_kind && !sym_fits(K, cfg, grad)) {
            while (sub > 1 && !sym_fits(sub, cfg, grad)) sub = (sub + 1) / 2;
            if (!sym_fits(sub, cfg, grad)) {
                if (kernel_ == HPH_PAIR_KERNEL_SYMMETRIC || kernel_ == HPH_PAIR_KERNEL_SPLIT)
                    throw std::runtime_error("the symmetric pass's partial sums do not fit in device memory at "
                                             "this N; choose the row-wise pass");
                cfg = TuneConfig{}; // row-wise: partials of K x chunks x 6 x N doubles only
                sub = K;
            }
        }
        double *w_dev = nullptr; // 1 / lambda of vector 0 (sorted order), kept for the location gradient
        if (write_probs) {
            d_w_.ensure(sizeof(double) * npad_);
            w_dev = d_w_.as<double>();
        }
        if (cfg.kind == HPH_PAIR_KERNEL_SYMMETRIC || cfg.kind == HPH_PAIR_KERNEL_SPLIT) {
            SymRunner::DataRef ref{ev.x, ev.t, ev.perm, ev.n, ev.npad, ev.t_last};
            const bool split = cfg.kind == HPH_PAIR_KERNEL_SPLIT ||
                               (kernel_ == HPH_PAIR_KERNEL_AUTO && skip_zero && split_pays(params, K));
            last_fast_ = !force_natural_ && fast_ok(extents_, params, K, split);
            last_ktab_ = last_fast_ && !split && !fp32_ && ktab_ok(extents_, params, K);
            if (split && !culled_valid_) build_culled();
            // vectors [k0, k0 + kk) per launch (one launch unless the partial sums of all K do not fit;
            // probabilities, 1/lambda and the cached sums belong to single-vector calls, never split);
            // with timing, the first launch's start and the last launch's end are recorded
            for (int k0 = 0; k0 < K; k0 += sub) {
                const int kk = std::min(sub, K - k0);
                const bool first = k0 == 0, last = k0 + kk == K;
                hipEvent_t e0 = timing_ && first ? ev_[0] : nullptr;
                hipEvent_t e1 = timing_ && last ? ev_[1] : nullptr, e2 = timing_ && last ? ev_[2] : nullptr;
                const double *pdev = d_params_.as<double>() + 6 * (size_t)k0;
                const double *phost = h_params_.as<double>() + 6 * (size_t)k0;
                double *rdev = d_result_.as<double>() + (size_t)kern::kResultWidth * k0;
                if (split) {
                    d_bsum_.ensure(sizeof(double) * (size_t)kk * (grad ? 3 : 1) * npad_);
                    d_sband_.ensure(sizeof(double) * (size_t)kk * 3 * npad_);
                    SymRunner::DataRef culled{d_xc_.as<double>(), d_tc_.as<double>(), d_cperm_.as<int>(), ev.n,
                                              ev.npad, ev.t_last};
                    sym_.enqueue_split(ref, culled, d_gstat_.as<double>(), kk, embeddingDimension, pdev, grad, fp32_,
                                       cfg.rows_per_lane, cfg.target_blocks, skip_zero, exact_s_ ? 0 : 1, last_fast_,
                                       d_bsum_.as<double>(), d_sband_.as<double>(), probs_dev, w_dev, rdev, stream_,
                                       e0, e1, e2, sums, gstat_host_.data(), phost);
                } else {
                    sym_.enqueue({ref}, {kk}, embeddingDimension, pdev, grad, fp32_, cfg.rows_per_lane,
                                 cfg.target_blocks, skip_zero, exact_s_ ? 0 : 1,
                                 {last_ktab_ ? (char)2 : last_fast_ ? (char)1 : (char)0},
                                 probs_dev, w_dev, rdev, stream_, e0, e1, e2, sums);
                }
            }
            last_sub_batch_ = sub;
            last_k_ = K;
            last_split_ = split;
            last_kind_ = split ? HPH_PAIR_KERNEL_SPLIT : HPH_PAIR_KERNEL_SYMMETRIC;
        } else {
            last_fast_ = false;
            last_ktab_ = false;
            last_split_ = false;
            last_kind_ = HPH_PAIR_KERNEL_ROWWISE;
            last_k_ = K;
            kern::PairGrid g = kern::plan_pair_grid(locationCount, K, cfg.target_blocks);
            g.skip_zero = skip_zero;
            g.zero_arg_b = kern::background_skip_arg(locationCount, fp32_);
            g.zero_arg_s = kern::kZeroExpArg;
            g.cut_s = skip_zero && !exact_s_ ? 1 : 0;
            const int nacc = grad ? 6 : 2;
            d_part_.ensure(sizeof(double) * (size_t)K * g.chunks * nacc * npad_);
            if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            kern::launch_pair(ev, d_params_.as<double>(), K, g, grad, fp32_, d_part_.as<double>(), stream_);
            check(hipGetLastError(), "pair kernel launch");
            if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
            kern::launch_finalize(ev, d_params_.as<double>(), K, g, grad, d_part_.as<double>(),
                                  d_block_.as<double>(), probs_dev, w_dev, sums, stream_);
            check(hipGetLastError(), "finalize kernel launch");
            kern::launch_reduce(ev, K, fin_blocks, grad, d_block_.as<double>(), d_result_.as<double>(), stream_);
            check(hipGetLastError(), "reduce kernel launch");
            if (timing_) check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
        }
        check(hipMemcpyAsync(h_result_.ptr, d_result_.ptr, sizeof(double) * (size_t)K * kern::kResultWidth,
                             hipMemcpyDeviceToHost, stream_),
              "download result");
        if (probs_host)
            check(hipMemcpyAsync(probs_host, d_probs_.ptr, sizeof(double) * locationCount, hipMemcpyDeviceToHost,
                                 stream_),
                  "download probs");
        check(hipStreamSynchronize(stream_), "stream sync");
        if (timing_) record_timing();
    }

    // Device times of the last evaluation: [0] = ev_[0] .. ev_[1], [1] = ev_[1] .. ev_[2].
    void record_timing() {
        float a = 0.f, b = 0.f;
        check(hipEventElapsedTime(&a, ev_[0], ev_[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&b, ev_[1], ev_[2]), "hipEventElapsedTime");
        last_ms_[0] = a;
        last_ms_[1] = b;
    }

    // Re-evaluation from the per-event moments of the last full single-vector evaluation (the
    // moments the reference's random-scan sampler leaves unchanged, R/Untitled.R:254-422, which
    // proposes sigma, omega, theta and mu0 only): with tauX and tauT unchanged the background
    // moments are reused; with sigma and omega unchanged as well the excitation moments are reused
    // too (theta and mu0 enter only as coefficients), else an excitation pass forms them. Reused
    // excitation moments were cut at the skip point of their own vector; they serve while the
    // current vector's point is at most ln 2 above it (the cut then moves every rate by under
    // 4e-22 relative). Dimensions 1..3; returns false when a full evaluation is needed.
    bool evaluate_incremental(bool grad) {
        if (!cache_valid_ || data_dirty_ || embeddingDimension > 3) return false;
        if (grad && !cache_grad_) return false;
        if (params_[1] != cache_params_[1] || params_[2] != cache_params_[2]) return false;
        const double xs = kern::excitation_skip_arg(params_, embeddingDimension, locationCount);
        const bool reuse_s = params_[0] == cache_params_[0] && params_[3] == cache_params_[3] &&
                             xs <= cache_xs_ + 0.69314718055994530942;
        TuneConfig cfg = config_for(1, grad);
        if (cfg.kind != HPH_PAIR_KERNEL_SYMMETRIC && cfg.kind != HPH_PAIR_KERNEL_SPLIT) {
            cfg.kind = HPH_PAIR_KERNEL_SYMMETRIC;
            cfg.rows_per_lane = locationCount < kSymA1MaxEvents ? 1 : 2;
            cfg.target_blocks = 4096;
        }
        if (!sym_fits(1, cfg, grad)) return false;
        check(hipSetDevice(device_), "hipSetDevice");
        d_params_.ensure(sizeof(double) * 6);
        const int fin_blocks = (locationCount + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock;
        d_block_.ensure(sizeof(double) * (size_t)fin_blocks * kern::kResultWidth * 2);
        d_result_.ensure(sizeof(double) * kern::kResultWidth);
        h_result_.ensure(sizeof(double) * kern::kResultWidth);
        h_params_.ensure(sizeof(double) * 6);
        d_w_.ensure(sizeof(double) * npad_);
        std::memcpy(h_params_.ptr, params_, sizeof(params_));
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6, hipMemcpyHostToDevice, stream_),
              "upload params");
        const int nxt = 1 - sums_cur_;
        d_sums_[nxt].ensure(sizeof(double) * 6 * (size_t)npad_);
        const kern::Events ev = events();
        SymRunner::DataRef ref{ev.x, ev.t, ev.perm, ev.n, ev.npad, ev.t_last};
        last_fast_ = false;
        last_ktab_ = false;
        const double *cached = d_sums_[sums_cur_].as<double>();
        const double *scache = cached + 3 * (size_t)npad_;
        if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
        if (!reuse_s) {
            d_sband_.ensure(sizeof(double) * 3 * (size_t)npad_);
            kern::launch_excitation_band(ev, d_params_.as<double>(), 1, grad, 0, d_sband_.as<double>(), stream_);
            check(hipGetLastError(), "excitation pass launch");
            scache = d_sband_.as<double>();
        }
        if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
        sym_.enqueue_from_sums(ref, embeddingDimension, d_params_.as<double>(), grad, fp32_, cfg.rows_per_lane,
                               cfg.target_blocks, cached, scache, d_sums_[nxt].as<double>(), d_probs_.as<double>(),
                               d_w_.as<double>(), d_result_.as<double>(), stream_, timing_ ? ev_[2] : nullptr);
        check(hipMemcpyAsync(h_result_.ptr, d_result_.ptr, sizeof(double) * kern::kResultWidth, hipMemcpyDeviceToHost,
                             stream_),
              "download result");
        check(hipStreamSynchronize(stream_), "stream sync");
        if (timing_) record_timing();
        sums_cur_ = nxt;
        last_eval_ = reuse_s ? HPH_EVAL_CACHED : HPH_EVAL_EXCITATION;
        if (!grad) cache_grad_ = false; // only moments 0 and 3 were written
        std::memcpy(cache_params_, params_, sizeof(params_));
        if (!reuse_s) cache_xs_ = xs;
        return true;
    }

    // A probability request at the tauX and tauT of the cached per-event sums (kept by the last
    // full or re-evaluated log-likelihood call): their background moments are the numbers a
    // probability pass would form (both passes cut the background at the same point), so only the
    // excitation moments are formed anew, by the band kernel skipping exact zeros alone (the
    // probability passes' cut), then the finalize from the sums writes the probabilities. The
    // cached sums, the cached log-likelihood and 1/lambda are left as they were (a later
    // log-likelihood call re-evaluates from the same sums as before). Dimensions 1..3, not with
    // HPH_FLAG_ALL_PAIRS; returns false when no such sums are at hand (the caller then runs the
    // full probability pass).
    bool probs_from_cache(double *probs_host) {
        if (!have_params_ || !cache_valid_ || data_dirty_ || embeddingDimension > 3 || (flags & ALL_PAIRS) ||
            force_natural_)
            return false;
        if (params_[1] != cache_params_[1] || params_[2] != cache_params_[2]) return false;
        const bool plan_grad = cache_grad_; // the plan of the cached evaluation: no re-planning
        TuneConfig cfg = config_for(1, plan_grad);
        if (cfg.kind != HPH_PAIR_KERNEL_SYMMETRIC && cfg.kind != HPH_PAIR_KERNEL_SPLIT) {
            cfg.kind = HPH_PAIR_KERNEL_SYMMETRIC;
            cfg.rows_per_lane = locationCount < kSymA1MaxEvents ? 1 : 2;
            cfg.target_blocks = 4096;
        }
        if (!sym_fits(1, cfg, plan_grad)) return false;
        check(hipSetDevice(device_), "hipSetDevice");
        d_params_.ensure(sizeof(double) * 6);
        const int fin_blocks = (locationCount + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock;
        d_block_.ensure(sizeof(double) * (size_t)fin_blocks * kern::kResultWidth * 2);
        d_result_.ensure(sizeof(double) * kern::kResultWidth);
        h_params_.ensure(sizeof(double) * 6);
        d_probs_.ensure(sizeof(double) * (size_t)locationCount);
        d_sband_.ensure(sizeof(double) * 3 * (size_t)npad_);
        d_sums_[1 - sums_cur_].ensure(sizeof(double) * 6 * (size_t)npad_);
        std::memcpy(h_params_.ptr, params_, sizeof(params_));
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6, hipMemcpyHostToDevice, stream_),
              "upload params");
        const kern::Events ev = events();
        SymRunner::DataRef ref{ev.x, ev.t, ev.perm, ev.n, ev.npad, ev.t_last};
        if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
        // probabilities need the sums S and B only: the band kernel and the finalize without moments
        kern::launch_excitation_band(ev, d_params_.as<double>(), 1, false, 3, d_sband_.as<double>(), stream_);
        check(hipGetLastError(), "excitation pass launch");
        if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
        // the moments written go to the spare sums buffer, which the cache does not point at
        sym_.enqueue_from_sums(ref, embeddingDimension, d_params_.as<double>(), false, fp32_, cfg.rows_per_lane,
                               cfg.target_blocks, d_sums_[sums_cur_].as<double>(), d_sband_.as<double>(),
                               d_sums_[1 - sums_cur_].as<double>(), d_probs_.as<double>(), nullptr,
                               d_result_.as<double>(), stream_, timing_ ? ev_[2] : nullptr, plan_grad ? 1 : 0);
        check(hipMemcpyAsync(probs_host, d_probs_.ptr, sizeof(double) * locationCount, hipMemcpyDeviceToHost, stream_),
              "download probs");
        check(hipStreamSynchronize(stream_), "stream sync");
        if (timing_) record_timing();
        probs_valid_ = true;
        last_eval_ = HPH_EVAL_EXCITATION;
        return true;
    }

    // A batch whose K vectors share tauX and tauT (the reference sampler's random scan proposes
    // sigma, omega, theta and mu0, R/Untitled.R:254-422; multiple-try proposals and parallel
    // chains of that scan share them too): the background moments depend on the data, tauX and
    // tauT alone, so they are formed once -- taken from the single-vector cache when it holds them
    // for these tauX and tauT, else by a full single-vector pass for vector 0, which refills the
    // cache -- and each vector gets its excitation moments from the band kernel (its own
    // likelihood-pass cut X_S) and a finalize. Same results as K full passes to rounding.
    // Dimensions 1..3, at least two vectors, not with HPH_FLAG_ALL_PAIRS; returns false when the
    // batch does not qualify (the caller then runs the K-vector pass).
    bool evaluate_batch_shared(int K, bool grad) {
        if (K < 2 || embeddingDimension > 3 || (flags & ALL_PAIRS)) return false;
        const double *P = batch_.data();
        for (int k = 1; k < K; ++k)
            if (P[6 * (size_t)k + 1] != P[1] || P[6 * (size_t)k + 2] != P[2]) return false;
        require_ready();
        check(hipSetDevice(device_), "hipSetDevice");
        if (data_dirty_) upload();
        const int N = locationCount, D = embeddingDimension;
        const bool have = cache_valid_ && (!grad || cache_grad_) && cache_params_[1] == P[1] && cache_params_[2] == P[2];
        if (!have) {
            // a full pass for vector 0 whose per-event moments become the single-vector cache
            d_sums_[sums_cur_].ensure(sizeof(double) * 6 * (size_t)npad_);
            cache_valid_ = false;
            run(P, 1, grad, nullptr, false, d_sums_[sums_cur_].as<double>());
            cache_valid_ = true;
            cache_grad_ = grad;
            std::memcpy(cache_params_, P, sizeof(cache_params_));
            cache_xs_ = kern::excitation_skip_arg(P, D, N);
        }
        const int NQ = grad ? 3 : 1;
        d_params_.ensure(sizeof(double) * 6 * (size_t)K);
        d_result_.ensure(sizeof(double) * (size_t)K * kern::kResultWidth);
        h_result_.ensure(sizeof(double) * (size_t)K * kern::kResultWidth);
        h_params_.ensure(sizeof(double) * 6 * (size_t)K);
        d_sband_batch_.ensure(sizeof(double) * (size_t)K * NQ * npad_);
        std::memcpy(h_params_.ptr, P, sizeof(double) * 6 * (size_t)K);
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6 * (size_t)K, hipMemcpyHostToDevice,
                             stream_),
              "upload params");
        const kern::Events ev = events();
        if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
        kern::launch_excitation_band(ev, d_params_.as<double>(), K, grad, 0, d_sband_batch_.as<double>(), stream_);
        check(hipGetLastError(), "excitation pass launch");
        SymRunner::DataRef ref{ev.x, ev.t, ev.perm, ev.n, ev.npad, ev.t_last};
        sym_batch_.enqueue_shared_background(ref, D, d_params_.as<double>(), grad, fp32_, K,
                                             d_sums_[sums_cur_].as<double>(), d_sband_batch_.as<double>(),
                                             d_result_.as<double>(), stream_, timing_ ? ev_[1] : nullptr,
                                             timing_ ? ev_[2] : nullptr);
        check(hipMemcpyAsync(h_result_.ptr, d_result_.ptr, sizeof(double) * (size_t)K * kern::kResultWidth,
                             hipMemcpyDeviceToHost, stream_),
              "download result");
        check(hipStreamSynchronize(stream_), "stream sync");
        if (timing_) record_timing();
        return true;
    }

    // Single-vector evaluation; caches ll (and the gradient). The probabilities it writes are kept
    // only when the pass skipped exact zeros alone: a request for probabilities runs such a pass
    // (excitation_skip_arg bounds the cut absolutely, not relatively, for tiny probabilities).
    void evaluate_single(bool grad, double *probs_host) {
        if (!have_params_) throw NotInitialized("parameters are not set (setParameters)");
        exact_s_ = probs_host != nullptr || (flags & ALL_PAIRS);
        bool incremental = false;
        try {
            incremental = !exact_s_ && !force_natural_ && evaluate_incremental(grad);
            if (!incremental) {
                // a full evaluation; its per-event moments are kept for later ones
                const bool keep = embeddingDimension <= 3 && !exact_s_;
                double *sums = nullptr;
                if (keep) {
                    d_sums_[sums_cur_].ensure(sizeof(double) * 6 * (size_t)npad_);
                    sums = d_sums_[sums_cur_].as<double>();
                }
                cache_valid_ = false;
                run(params_, 1, grad, probs_host, true, sums);
                last_eval_ = HPH_EVAL_FULL;
                if (keep) {
                    cache_valid_ = true;
                    cache_grad_ = grad;
                    std::memcpy(cache_params_, params_, sizeof(params_));
                    cache_xs_ = kern::excitation_skip_arg(params_, embeddingDimension, locationCount);
                }
            }
        } catch (...) {
            exact_s_ = false;
            throw;
        }
        probs_valid_ = exact_s_;
        exact_s_ = false;
        // the location gradient takes 1/lambda from natural-form FP64 full passes only: its
        // components cancel between neighbouring events, so it needs rates to about 1e-15
        w_valid_ = !fp32_ && !last_fast_ && !incremental;
        const double *r = h_result_.as<double>();
        ll_ = r[0];
        ll_valid_ = true;
        if (grad) {
            std::memcpy(grad_, r + 1, sizeof(grad_));
            grad_valid_ = true;
        }
    }

    // Built-in choice without tuning (one MI355X, LL + gradient, K = 1): row-wise narrow below
    // 12,288 events (41 us at N = 1e3, 70 us at 5e3, 99 us at 8e3), symmetric with 1 row per lane
    // below 24,576 (229 us at 1.6e4), symmetric with 2 rows per lane above (551 us at 3.3e4).
    static constexpr int kNarrowMaxEvents = 12288;
    static constexpr int kSymA1MaxEvents = 24576;
    int kernel_ = HPH_PAIR_KERNEL_AUTO;
    SymRunner sym_;
    Extents extents_;
    std::vector<double> gstat_host_; // the culled order's box table (host), for its queue order
    int last_sub_batch_ = 0;         // vectors per launch of the last symmetric / split evaluation
    int last_k_ = 0;                 // vectors of the last full evaluation

    int device_;
    bool fp32_;
    hipStream_t stream_ = nullptr;
    int npad_ = 0;

    std::vector<double> locs_, times_;
    std::vector<int> perm_;
    bool have_locs_ = false, have_times_ = false, have_params_ = false;
    bool data_dirty_ = true, order_dirty_ = true;
    double params_[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> batch_;

    double ll_ = 0.0, grad_[6] = {0, 0, 0, 0, 0, 0};
    bool ll_valid_ = false, grad_valid_ = false, probs_valid_ = false, w_valid_ = false;
    bool force_natural_ = false, last_fast_ = false, last_ktab_ = false, exact_s_ = false, last_split_ = false;
    int last_kind_ = 0; // pass of the last full evaluation (HPH_PAIR_KERNEL_*), 0 before the first
    std::map<long long, bool> sym_fits_;

    bool has_stored_ = false, stored_have_params_ = false, stored_have_locs_ = false, stored_ll_valid_ = false;
    double stored_params_[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> stored_locs_;
    double stored_ll_ = 0.0;

    TuneConfig tuned_, forced_;
    bool timing_ = false;
    hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};
    double last_ms_[2] = {0.0, 0.0};

    DevBuf d_x_, d_t_, d_perm_, d_probs_, d_params_, d_part_, d_block_, d_result_, d_w_, d_locgrad_;
    HostBuf h_stage_, h_result_, h_params_;
    // culled order of the split passes (built on first use after a data change)
    bool culled_valid_ = false;
    DevBuf d_xc_, d_tc_, d_cperm_, d_gstat_, d_bsum_;
    // per-event moments of the last single-vector evaluation (two buffers: one read, one written)
    DevBuf d_sums_[2];
    DevBuf d_sband_; // excitation sums of a re-evaluation ([3][npad])
    DevBuf d_sband_batch_; // excitation sums of a batch sharing tauX and tauT ([K][3 or 1][npad])
    SymRunner sym_batch_;  // finalize plans of such batches (kept apart from the single-vector plan)
    int sums_cur_ = 0;
    bool cache_valid_ = false, cache_grad_ = false;
    double cache_params_[6] = {0, 0, 0, 0, 0, 0};
    double cache_xs_ = 0.0; // excitation skip point the cached excitation moments were cut at
    int last_eval_ = HPH_EVAL_FULL;
};


// Many datasets of uneven size, each with its own parameter vectors, evaluated by one symmetric
// pass (one persistent launch over a cost-ordered work queue). With graph mode on, the launch
// sequence (parameter upload, queue reset, pass, finalize, reduction, result download) is
// captured once per shape and replayed.
class HipHawkesBatch {
  public:
    HipHawkesBatch(int D, const std::vector<int> &counts, long flags, int device)
        : D_(D), counts_(counts), flags_(flags), device_(device < 0 ? 0 : device), fp32_((flags & FLOAT) != 0) {
        if (D < 1 || D > 3) throw std::invalid_argument("batched datasets support dimensions 1..3");
        if (counts.empty()) throw std::invalid_argument("dataset_count must be >= 1");
        for (int c : counts) {
            if (c < 1) throw std::invalid_argument("every dataset needs at least one event");
            if (c > (1 << 30)) throw std::invalid_argument("a dataset has at most 2^30 events");
        }
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device_ >= count) throw std::invalid_argument("device index out of range");
        check(hipSetDevice(device_), "hipSetDevice");
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        const int nd = (int)counts.size();
        off_.resize(nd + 1, 0);
        npad_.resize(nd);
        for (int d = 0; d < nd; ++d) {
            npad_[d] = ((counts[d] + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock) * kern::kRowsPerBlock;
            off_[d + 1] = off_[d] + npad_[d];
        }
        locs_.resize(nd);
        times_.resize(nd);
        sorted_.resize(nd);
        split_.resize(nd);
        ext_.resize(nd);
        have_.assign(nd, false);
        d_x_.ensure(sizeof(double) * (size_t)D * off_[nd]);
        d_t_.ensure(sizeof(double) * (size_t)off_[nd]);
        d_perm_.ensure(sizeof(int) * (size_t)off_[nd]);
        // timing only (read after a stream synchronisation): no system-scope fence
        for (auto &e : ev_) check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
    }

    ~HipHawkesBatch() {
        (void)hipSetDevice(device_);
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graph();
        for (DevBuf *b : {&d_x_, &d_t_, &d_perm_, &d_params_, &d_result_}) b->release();
        for (auto &sd : split_)
            if (sd)
                for (DevBuf *b : {&sd->x, &sd->t, &sd->perm, &sd->gstat, &sd->bsum, &sd->sband}) b->release();
        split_.clear();
        h_params_.release();
        h_result_.release();
        for (auto &e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    void setData(int d, const double *locs, size_t locs_len, const double *times, size_t times_len) {
        if (d < 0 || d >= (int)counts_.size()) throw std::invalid_argument("dataset index out of range");
        const int n = counts_[d];
        if (locs == nullptr || times == nullptr) throw std::invalid_argument("data pointer is null");
        if (locs_len != (size_t)n * D_ || times_len != (size_t)n)
            throw std::invalid_argument("dataset lengths must be N*D locations and N times");
        for (size_t q = 0; q < locs_len; ++q)
            if (!std::isfinite(locs[q])) throw std::invalid_argument("locations must be finite");
        for (size_t q = 0; q < times_len; ++q)
            if (!std::isfinite(times[q])) throw std::invalid_argument("times must be finite");
        check(hipSetDevice(device_), "hipSetDevice");
        locs_[d].assign(locs, locs + locs_len);
        times_[d].assign(times, times + times_len);
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return times[a] < times[b]; });
        const int np = npad_[d];
        std::vector<double> hx((size_t)D_ * np), ht(np);
        std::vector<int> hp(np);
        for (int s = 0; s < np; ++s) {
            const int u = perm[s < n ? s : n - 1];
            for (int q = 0; q < D_; ++q) hx[(size_t)q * np + s] = locs[(size_t)u * D_ + q];
            ht[s] = times[u];
            hp[s] = u;
        }
        ext_[d] = extents_of(locs_[d], ht.data(), (size_t)n, D_);
        sorted_[d] = perm;
        if (split_[d]) split_[d]->valid = false;
        // x planes of dataset d live at d_x_ + D * off_[d]: plane q at + q * npad_[d]
        check(hipMemcpy(d_x_.as<double>() + (size_t)D_ * off_[d], hx.data(), sizeof(double) * hx.size(),
                        hipMemcpyHostToDevice),
              "upload x");
        check(hipMemcpy(d_t_.as<double>() + off_[d], ht.data(), sizeof(double) * np, hipMemcpyHostToDevice),
              "upload t");
        check(hipMemcpy(d_perm_.as<int>() + off_[d], hp.data(), sizeof(int) * np, hipMemcpyHostToDevice),
              "upload perm");
        have_[d] = true;
        drop_graph();
    }

    void setParameters(const int *vectors, const double *params, size_t count) {
        if (vectors == nullptr || params == nullptr) throw std::invalid_argument("parameter pointer is null");
        const int nd = (int)counts_.size();
        std::vector<int> K(vectors, vectors + nd);
        size_t total = 0;
        for (int k : K) {
            if (k < 1) throw std::invalid_argument("every dataset needs at least one parameter vector");
            total += (size_t)k;
        }
        if (count != total * 6) throw std::invalid_argument("count must be 6 x the total number of vectors");
        for (size_t q = 0; q < count; ++q)
            if (!std::isfinite(params[q])) throw std::invalid_argument("parameters must be finite");
        if (K != K_) drop_graph();
        K_ = K;
        total_ = (int)total;
        params_.assign(params, params + count);
        have_params_ = true;
    }

    void setGraph(bool enabled) {
        if (!enabled) drop_graph();
        graph_mode_ = enabled;
    }

    // A fixed work decomposition (rows per lane, target number of work items) instead of the tuned
    // or built-in one.
    void setConfig(int rows_per_lane, int target_items) {
        if (rows_per_lane != 1 && rows_per_lane != 2) throw std::invalid_argument("rows per lane must be 1 or 2");
        if (target_items < 1 || target_items > (1 << 24)) throw std::invalid_argument("target items out of range");
        forced_ = true;
        forced_A_ = rows_per_lane;
        forced_target_ = target_items;
    }

    // Times each work decomposition (1 or 2 rows per lane; six target item counts) on the current
    // datasets and vectors (eager launches, device time, the best of `repeats`), keeps the fastest
    // for this table's shape in the process-wide tuning cache (and the HPH_TUNE_CACHE file), and
    // uses it from then on unless a configuration is set explicitly.
    void autotune(int repeats, double *best_ms) {
        if (!have_params_) throw NotInitialized("parameters are not set");
        if (repeats < 1) repeats = 1;
        std::vector<double> ll((size_t)total_), g((size_t)total_ * 6);
        const bool was_graph = graph_mode_, was_forced = forced_;
        graph_mode_ = false;
        double best_t = 1e300;
        TuneConfig best;
        best.kind = HPH_PAIR_KERNEL_SYMMETRIC;
        try {
            for (int a : {1, 2}) {
                for (int target : {2048, 4096, 8192, 16384, 32768, 65536}) {
                    forced_ = true;
                    forced_A_ = a;
                    forced_target_ = target;
                    evaluate(ll.data(), g.data(), ll.size()); // warm: plans, buffers
                    double t = 1e300;
                    for (int r = 0; r < repeats; ++r) {
                        evaluate(ll.data(), g.data(), ll.size());
                        t = std::min(t, last_ms_);
                    }
                    if (t < best_t) {
                        best_t = t;
                        best.rows_per_lane = a;
                        best.target_blocks = target;
                    }
                }
            }
        } catch (...) {
            graph_mode_ = was_graph;
            forced_ = was_forced;
            throw;
        }
        graph_mode_ = was_graph;
        forced_ = false;
        tune_cache().store(tune_key(true), best);
        tune_cache().store(tune_key(false), best);
        if (best_ms) *best_ms = best_t;
    }

    // The work decomposition the next evaluation uses: out[0] rows per lane, out[1] target items,
    // out[2] = 0 built-in, 1 tuned, 2 set explicitly.
    void getConfig(int *out, size_t length) {
        if (out == nullptr || length < 3) throw std::invalid_argument("config output needs 3 values");
        int source = 0;
        const TuneConfig c = effective_config(grad_, &source);
        out[0] = c.rows_per_lane;
        out[1] = c.target_blocks;
        out[2] = source;
    }

    void evaluate(double *loglik, double *grad, size_t total) {
        if (loglik == nullptr) throw std::invalid_argument("loglik pointer is null");
        for (size_t d = 0; d < have_.size(); ++d)
            if (!have_[d]) throw NotInitialized("dataset " + std::to_string(d) + " has no data (set_data)");
        if (!have_params_) throw NotInitialized("parameters are not set");
        if (total != (size_t)total_) throw std::invalid_argument("total does not match the parameter vectors");
        check(hipSetDevice(device_), "hipSetDevice");
        const bool want_grad = grad != nullptr;
        if (want_grad != grad_) drop_graph();
        grad_ = want_grad;
        // the work decomposition: set explicitly, tuned for this shape, or built in; the graph
        // holds one
        const TuneConfig cfg = effective_config(grad_, nullptr);
        if (cfg.rows_per_lane != A_ || cfg.target_blocks != target_) {
            A_ = cfg.rows_per_lane;
            target_ = cfg.target_blocks;
            drop_graph();
        }
        d_params_.ensure(sizeof(double) * 6 * (size_t)total_);
        d_result_.ensure(sizeof(double) * kern::kResultWidth * (size_t)total_);
        h_result_.ensure(sizeof(double) * kern::kResultWidth * (size_t)total_);
        h_params_.ensure(sizeof(double) * 6 * (size_t)total_);
        const std::vector<char> split = split_flags();
        const std::vector<char> fast = fast_flags(split);
        // the graph holds one variant, table and split layout
        if (graph_exec_ && (fast != graph_fast_ || split != graph_split_)) drop_graph();
        // vector slots: the datasets of the time-sorted pass first (dataset order), then each
        // dataset of the split passes; parameters are staged and results read in that order
        const int nd = (int)counts_.size();
        slot_.assign(nd, 0);
        std::vector<int> first(nd, 0);
        for (int d = 1; d < nd; ++d) first[d] = first[d - 1] + K_[d - 1];
        int next = 0;
        for (int pass = 0; pass < 2; ++pass)
            for (int d = 0; d < nd; ++d)
                if ((split[d] != 0) == (pass == 1)) {
                    slot_[d] = next;
                    next += K_[d];
                }
        for (int d = 0; d < nd; ++d)
            std::memcpy(h_params_.as<double>() + 6 * (size_t)slot_[d], params_.data() + 6 * (size_t)first[d],
                        sizeof(double) * 6 * (size_t)K_[d]);
        for (int d = 0; d < nd; ++d)
            if (split[d]) ensure_split(d); // host work and uploads, outside any capture
        if (graph_mode_) {
            if (!graph_exec_) {
                enqueue(false, fast, split); // plans, allocates and uploads the tables outside the capture
                check(hipStreamSynchronize(stream_), "stream sync");
                check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                enqueue(false, fast, split);
                graph_fast_ = fast;
                graph_split_ = split;
                hipGraph_t g = nullptr;
                check(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
                graph_ = g;
                check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
            }
            // the culled passes' work-queue order for these parameters, into the staging buffers
            // the graph copies (the time-sorted pass orders its queue on the device)
            if (!(flags_ & ALL_PAIRS))
                for (int d = 0; d < nd; ++d)
                    if (split[d])
                        split_[d]->runner.order_queue_culled(split_[d]->gstat_host.data(),
                                                             h_params_.as<double>() + 6 * (size_t)slot_[d], D_);
            check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
            check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
        } else {
            enqueue(true, fast, split);
        }
        check(hipStreamSynchronize(stream_), "stream sync");
        float a = 0.f;
        check(hipEventElapsedTime(&a, ev_[0], ev_[2]), "hipEventElapsedTime");
        last_ms_ = a;
        last_split_ = split;
        const double *r = h_result_.as<double>();
        for (int d = 0; d < nd; ++d)
            for (int k = 0; k < K_[d]; ++k) {
                const double *rk = r + (size_t)(slot_[d] + k) * kern::kResultWidth;
                const size_t o = (size_t)first[d] + k;
                loglik[o] = rk[0];
                if (want_grad)
                    for (int q = 0; q < 6; ++q) grad[o * 6 + q] = rk[1 + q];
            }
    }

    // Whether dataset d took the split passes in the last evaluation.
    bool lastSplit(int d) const {
        if (d < 0 || d >= (int)counts_.size()) throw std::invalid_argument("dataset index out of range");
        return d < (int)last_split_.size() && last_split_[d] != 0;
    }

    double lastMs() const { return last_ms_; }

  private:
    // The tuning-cache key of this table: log2 of the pair-evaluations of a launch, dimension,
    // total vectors, gradient, precision, dataset count.
    TuneKey tune_key(bool grad) const {
        double pairs = 0.0;
        for (size_t d = 0; d < counts_.size(); ++d)
            pairs += (double)(d < K_.size() ? K_[d] : 1) * (double)counts_[d] * (double)counts_[d];
        TuneKey key;
        key.log2n = (int)std::floor(std::log2(std::max(pairs, 1.0)));
        key.dim = D_;
        key.K = total_;
        key.grad = grad ? 1 : 0;
        key.fp32 = fp32_ ? 1 : 0;
        key.datasets = (int)counts_.size();
        return key;
    }

    // source (may be null): 0 built-in (2 rows per lane, 8192 items), 1 tuned, 2 set explicitly.
    TuneConfig effective_config(bool grad, int *source) const {
        TuneConfig c;
        c.kind = HPH_PAIR_KERNEL_SYMMETRIC;
        c.rows_per_lane = 2;
        c.target_blocks = 8192;
        int src = 0;
        if (forced_) {
            c.rows_per_lane = forced_A_;
            c.target_blocks = forced_target_;
            src = 2;
        } else if (have_params_) {
            TuneConfig t;
            if (tune_cache().lookup(tune_key(grad), &t)) {
                c = t;
                src = 1;
            }
        }
        if (source) *source = src;
        return c;
    }

    // Per dataset: whether all its vectors of the current parameters allow the base-2 form (in the
    // split passes where split[d] is set): 1, or 2 where the time-sorted pass may also run
    // exp2_neg_ktab (FP64, ktab_ok).
    std::vector<char> fast_flags(const std::vector<char> &split) const {
        std::vector<char> fast(counts_.size());
        const double *hp = params_.data();
        for (size_t d = 0, k0 = 0; d < counts_.size(); k0 += (size_t)K_[d], ++d) {
            const bool f = fast_ok(ext_[d], hp + 6 * k0, K_[d], split[d] != 0);
            fast[d] = !f ? 0 : (!split[d] && !fp32_ && ktab_ok(ext_[d], hp + 6 * k0, K_[d])) ? 2 : 1;
        }
        return fast;
    }

    // Per dataset: whether it takes the split passes (likelihood passes with skipping, data spread
    // over several negligible-background distances for every one of its vectors; as for a single
    // engine's built-in choice).
    std::vector<char> split_flags() const {
        std::vector<char> split(counts_.size(), 0);
        if (flags_ & ALL_PAIRS) return split;
        const double *hp = params_.data();
        for (size_t d = 0, k0 = 0; d < counts_.size(); k0 += (size_t)K_[d], ++d)
            split[d] = split_pays(ext_[d], hp + 6 * k0, K_[d], counts_[d], D_, fp32_) ? 1 : 0;
        return split;
    }

    // The culled order of dataset d (make_culled) on the device, its runner and its per-event
    // sum buffers for the current vectors.
    void ensure_split(int d) {
        if (!split_[d]) split_[d] = std::make_unique<SplitData>();
        SplitData &sd = *split_[d];
        const int np = npad_[d];
        if (!sd.valid) {
            const CulledHost c = make_culled(locs_[d], sorted_[d], times_[d], counts_[d], D_, np);
            sd.x.ensure(sizeof(double) * c.x.size());
            sd.t.ensure(sizeof(double) * c.t.size());
            sd.perm.ensure(sizeof(int) * c.perm.size());
            sd.gstat.ensure(sizeof(double) * c.gstat.size());
            check(hipMemcpy(sd.x.ptr, c.x.data(), sizeof(double) * c.x.size(), hipMemcpyHostToDevice), "upload culled x");
            check(hipMemcpy(sd.t.ptr, c.t.data(), sizeof(double) * c.t.size(), hipMemcpyHostToDevice), "upload culled t");
            check(hipMemcpy(sd.perm.ptr, c.perm.data(), sizeof(int) * c.perm.size(), hipMemcpyHostToDevice),
                  "upload culled order");
            check(hipMemcpy(sd.gstat.ptr, c.gstat.data(), sizeof(double) * c.gstat.size(), hipMemcpyHostToDevice),
                  "upload group boxes");
            sd.gstat_host = c.gstat;
            sd.runner.reset_culled_order();
            sd.valid = true;
            drop_graph();
        }
        sd.bsum.ensure(sizeof(double) * (size_t)K_[d] * 3 * np);
        sd.sband.ensure(sizeof(double) * (size_t)K_[d] * 3 * np);
    }

    // Parameter upload, passes, finalize, reduction and result download on the batch's stream:
    // one time-sorted symmetric pass over the datasets not split, then each split dataset's
    // passes (background on its culled order, per-event sums, excitation band, finalize),
    // parameters and results in the slot order of evaluate.
    void enqueue(bool events, const std::vector<char> &fast, const std::vector<char> &split) {
        check(hipMemcpyAsync(d_params_.ptr, h_params_.ptr, sizeof(double) * 6 * (size_t)total_, hipMemcpyHostToDevice,
                             stream_),
              "upload params");
        const int skip_zero = (flags_ & ALL_PAIRS) ? 0 : 1;
        double *params = d_params_.as<double>(), *result = d_result_.as<double>();
        if (events) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
        std::vector<SymRunner::DataRef> refs;
        std::vector<int> K;
        std::vector<char> f;
        for (size_t d = 0; d < counts_.size(); ++d)
            if (!split[d]) {
                refs.push_back(ref(d));
                K.push_back(K_[d]);
                f.push_back(fast[d]);
            }
        if (!refs.empty()) {
            // the queue's order for this call's parameters, formed on the device from the uploaded
            // parameters (inside a captured graph too)
            runner_.enqueue(refs, K, D_, params, grad_, fp32_, A_, target_, skip_zero, 1, f, nullptr, nullptr, result,
                            stream_, nullptr, nullptr, nullptr, nullptr);
        }
        for (size_t d = 0; d < counts_.size(); ++d)
            if (split[d]) {
                SplitData &sd = *split_[d];
                const SymRunner::DataRef culled{sd.x.as<double>(), sd.t.as<double>(), sd.perm.as<int>(), counts_[d],
                                                npad_[d], times_[d].back()};
                sd.runner.enqueue_split(ref(d), culled, sd.gstat.as<double>(), K_[d], D_, params + 6 * (size_t)slot_[d],
                                        grad_, fp32_, A_, target_, skip_zero, 1, fast[d] != 0, sd.bsum.as<double>(),
                                        sd.sband.as<double>(), nullptr, nullptr,
                                        result + kern::kResultWidth * (size_t)slot_[d], stream_, nullptr, nullptr,
                                        nullptr, nullptr, sd.gstat_host.data(),
                                        h_params_.as<double>() + 6 * (size_t)slot_[d]);
            }
        if (events) check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
        check(hipMemcpyAsync(h_result_.ptr, d_result_.ptr, sizeof(double) * kern::kResultWidth * (size_t)total_,
                             hipMemcpyDeviceToHost, stream_),
              "download result");
    }

    SymRunner::DataRef ref(size_t d) const {
        return SymRunner::DataRef{d_x_.as<double>() + (size_t)D_ * off_[d], d_t_.as<double>() + off_[d],
                                  d_perm_.as<int>() + off_[d], counts_[d], npad_[d], times_[d].back()};
    }

    void drop_graph() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
    }

    int D_;
    std::vector<int> counts_;
    long flags_;
    int device_;
    bool fp32_;
    hipStream_t stream_ = nullptr;
    std::vector<long long> off_; // start of dataset d in the concatenated device arrays
    std::vector<int> npad_;
    std::vector<std::vector<double>> locs_, times_;
    std::vector<std::vector<int>> sorted_; // per dataset: time-sorted position -> caller's index
    std::vector<Extents> ext_;
    std::vector<bool> have_;
    std::vector<int> K_;
    std::vector<double> params_; // the caller's vectors, dataset order
    std::vector<int> slot_;      // result and parameter slot of dataset d's first vector
    int total_ = 0;
    bool have_params_ = false, grad_ = true, graph_mode_ = false;
    std::vector<char> graph_fast_, graph_split_, last_split_;
    // a dataset of the split passes: its culled order, runner and per-event sums
    struct SplitData {
        DevBuf x, t, perm, gstat, bsum, sband;
        std::vector<double> gstat_host; // the box table on the host (its queue order)
        bool valid = false;
        SymRunner runner;
    };
    std::vector<std::unique_ptr<SplitData>> split_;
    int A_ = 2, target_ = 8192; // work decomposition of the captured graph / last evaluation
    bool forced_ = false;
    int forced_A_ = 2, forced_target_ = 8192;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};
    double last_ms_ = 0.0;
    SymRunner runner_;
    DevBuf d_x_, d_t_, d_perm_, d_params_, d_result_;
    HostBuf h_params_, h_result_;
};

EngineSharedPtr engineFactory(int dim1, int dim2, long flags, int device, int /*threads*/) {
    return std::make_shared<HipHawkes>(dim1, dim2, flags, device);
}

SharedPtr factory(int dim1, int dim2, long flags, int device, int threads) {
    return engineFactory(dim1, dim2, flags, device, threads);
}

} // namespace hph

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

struct hph_engine {
    hph::EngineSharedPtr impl;
};

static thread_local std::string g_last_error;

template <class F> static int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return HPH_OK;
    } catch (const hph::NotInitialized &e) {
        g_last_error = e.what();
        return HPH_ERR_NOT_INITIALIZED;
    } catch (const std::invalid_argument &e) {
        g_last_error = e.what();
        return HPH_ERR_INVALID_ARGUMENT;
    } catch (const hph::DeviceOutOfMemory &e) {
        g_last_error = e.what();
        return HPH_ERR_OUT_OF_MEMORY;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return HPH_ERR_OUT_OF_MEMORY;
    } catch (const hph::DeviceError &e) {
        g_last_error = e.what();
        return HPH_ERR_DEVICE;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return HPH_ERR_DEVICE;
    } catch (...) {
        g_last_error = "unknown error";
        return HPH_ERR_DEVICE;
    }
}

#define HPH_REQUIRE_ENGINE(e)                                                                                  \
    if ((e) == nullptr || !(e)->impl) {                                                                        \
        g_last_error = "engine handle is null";                                                                \
        return HPH_ERR_INVALID_ARGUMENT;                                                                       \
    }

extern "C" {

const char *hph_last_error(void) { return g_last_error.c_str(); }

int hph_create(int embedding_dimension, int location_count, long flags, int device, int threads,
               hph_engine **out) {
    if (out == nullptr) {
        g_last_error = "output handle pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return guarded([&] {
        auto impl = hph::engineFactory(embedding_dimension, location_count, flags, device, threads);
        *out = new hph_engine{impl};
    });
}

int hph_destroy(hph_engine *engine) {
    if (engine == nullptr) return HPH_OK;
    return guarded([&] { delete engine; });
}

int hph_set_times_data(hph_engine *e, const double *data, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setTimesData(const_cast<double *>(data), length); });
}

int hph_update_locations(hph_engine *e, int location_index, const double *location, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->updateLocations(location_index, const_cast<double *>(location), length); });
}

int hph_set_parameters(hph_engine *e, const double *data, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setParameters(const_cast<double *>(data), length); });
}

int hph_get_sum_of_lik_contribs(hph_engine *e, double *out) {
    HPH_REQUIRE_ENGINE(e);
    if (out == nullptr) {
        g_last_error = "output pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *out = e->impl->getSumOfLikContribs(); });
}

int hph_get_log_like_gradient(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getLogLikeGradient(out, length); });
}

int hph_get_log_likelihood_and_gradient(hph_engine *e, double *loglik, double *grad) {
    HPH_REQUIRE_ENGINE(e);
    if (loglik == nullptr) {
        g_last_error = "output pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *loglik = e->impl->getLogLikelihoodAndGradient(grad); });
}

int hph_get_probs_self_excite(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getProbsSelfExcite(out, length); });
}

int hph_get_location_gradient(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getLocationGradient(out, length); });
}

int hph_store_state(hph_engine *e) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->storeState(); });
}

int hph_restore_state(hph_engine *e) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->restoreState(); });
}

int hph_accept_state(hph_engine *e) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->acceptState(); });
}

int hph_get_internal_dimension(hph_engine *e, int *out) {
    HPH_REQUIRE_ENGINE(e);
    if (out == nullptr) {
        g_last_error = "output pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *out = e->impl->getInternalDimension(); });
}

int hph_set_parameters_batch(hph_engine *e, const double *params, size_t count) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setParametersBatch(params, count); });
}

int hph_get_log_likelihood_batch(hph_engine *e, double *loglik, double *grad, size_t K) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getLogLikelihoodBatch(loglik, grad, K); });
}

int hph_set_timing(hph_engine *e, int enabled) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setTiming(enabled != 0); });
}

int hph_autotune(hph_engine *e, int repeats, double *best_ms) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->autotune(repeats, best_ms); });
}

int hph_get_tuned_config(hph_engine *e, int *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getTunedConfig(out, length); });
}

int hph_set_pair_config(hph_engine *e, int kind, int rows_per_lane, int target_blocks) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setPairConfig(kind, rows_per_lane, target_blocks); });
}

int hph_set_pair_kernel(hph_engine *e, int kind) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setPairKernel(kind); });
}

int hph_get_timing(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getTiming(out, length); });
}

int hph_get_last_pass(hph_engine *e, int *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getLastPass(out, length); });
}

int hph_get_event_sums(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getEventSums(out, length); });
}

int hph_get_last_evaluation(hph_engine *e, int *kind) {
    HPH_REQUIRE_ENGINE(e);
    if (kind == nullptr) {
        g_last_error = "kind pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *kind = e->impl->getLastEvaluation(); });
}

int hph_log_likelihood_noalloc(int embedding_dimension, int location_count, const double *locations,
                               const double *times, const double *params, int K, long flags, double *result,
                               int gradient, void *stream) {
    return guarded([&] {
        hph::noalloc::evaluate(embedding_dimension, location_count, locations, times, params, K, flags, result,
                               gradient != 0, static_cast<hipStream_t>(stream));
    });
}

struct hph_batch {
    hph::HipHawkesBatch *impl;
};

int hph_batch_create(int embedding_dimension, int dataset_count, const int *location_counts, long flags, int device,
                     hph_batch **out) {
    if (out == nullptr) {
        g_last_error = "output handle pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (dataset_count < 1 || location_counts == nullptr) {
        g_last_error = "dataset_count must be >= 1 and location_counts non-null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] {
        std::vector<int> counts(location_counts, location_counts + dataset_count);
        auto *impl = new hph::HipHawkesBatch(embedding_dimension, counts, flags, device);
        *out = new hph_batch{impl};
    });
}

int hph_batch_destroy(hph_batch *b) {
    if (b == nullptr) return HPH_OK;
    return guarded([&] {
        delete b->impl;
        delete b;
    });
}

#define HPH_REQUIRE_BATCH(b)                                                                                   \
    if ((b) == nullptr || (b)->impl == nullptr) {                                                              \
        g_last_error = "batch handle is null";                                                                 \
        return HPH_ERR_INVALID_ARGUMENT;                                                                       \
    }

int hph_batch_set_data(hph_batch *b, int dataset, const double *locations, size_t locations_length,
                       const double *times, size_t times_length) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->setData(dataset, locations, locations_length, times, times_length); });
}

int hph_batch_set_parameters(hph_batch *b, const int *vectors_per_dataset, const double *params, size_t count) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->setParameters(vectors_per_dataset, params, count); });
}

int hph_batch_get_log_likelihood(hph_batch *b, double *loglik, double *grad, size_t total_vectors) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->evaluate(loglik, grad, total_vectors); });
}

int hph_batch_set_graph(hph_batch *b, int enabled) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->setGraph(enabled != 0); });
}

int hph_batch_set_config(hph_batch *b, int rows_per_lane, int target_items) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->setConfig(rows_per_lane, target_items); });
}

int hph_batch_autotune(hph_batch *b, int repeats, double *best_ms) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->autotune(repeats, best_ms); });
}

int hph_batch_get_config(hph_batch *b, int *out, size_t length) {
    HPH_REQUIRE_BATCH(b);
    return guarded([&] { b->impl->getConfig(out, length); });
}

int hph_batch_get_timing(hph_batch *b, double *out) {
    HPH_REQUIRE_BATCH(b);
    if (out == nullptr) {
        g_last_error = "output pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *out = b->impl->lastMs(); });
}

int hph_batch_last_split(hph_batch *b, int dataset, int *out) {
    HPH_REQUIRE_BATCH(b);
    if (out == nullptr) {
        g_last_error = "output pointer is null";
        return HPH_ERR_INVALID_ARGUMENT;
    }
    return guarded([&] { *out = b->impl->lastSplit(dataset) ? 1 : 0; });
}

} // extern "C"
