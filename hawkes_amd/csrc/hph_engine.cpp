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
#include <cstring>
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

class HipHawkes : public AbstractHawkes {
  public:
    HipHawkes(int D, int N, long flags, int device)
        : AbstractHawkes(D, N, flags), device_(device < 0 ? 0 : device), fp32_((flags & FLOAT) != 0) {
        if (D < 1 || D > kern::kMaxDim) throw std::invalid_argument("embeddingDimension must be in 1..8");
        if (N < 1) throw std::invalid_argument("locationCount must be >= 1");
        if (fp32_) throw std::invalid_argument("FLOAT (FP32-bulk) mode is not built in this library yet");
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
        for (DevBuf *b : {&d_x_, &d_t_, &d_perm_, &d_probs_, &d_params_, &d_part_, &d_block_, &d_result_, &d_colpart_, &d_items_})
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

    void getProbsSelfExcite(double *result, size_t length) override {
        if (result == nullptr) throw std::invalid_argument("result pointer is null");
        if (length != (size_t)locationCount) throw std::invalid_argument("getProbsSelfExcite: length must be locationCount");
        evaluate_single(false, result);
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
        run(batch_.data(), (int)K, grad, nullptr);
        const double *r = h_result_.as<double>();
        for (size_t k = 0; k < K; ++k) {
            loglik[k] = r[k * kern::kResultWidth];
            if (grad)
                for (int q = 0; q < 6; ++q) gradient[k * 6 + q] = r[k * kern::kResultWidth + 1 + q];
        }
    }

    void setTiming(bool enabled) override {
        check(hipSetDevice(device_), "hipSetDevice");
        if (enabled && !ev_[0])
            for (auto &e : ev_) check(hipEventCreate(&e), "hipEventCreate");
        timing_ = enabled;
    }

    void setPairKernel(int kind) override {
        if (kind < HPH_PAIR_KERNEL_AUTO || kind > HPH_PAIR_KERNEL_SYMMETRIC)
            throw std::invalid_argument("unknown pair kernel kind");
        if (kind == HPH_PAIR_KERNEL_SYMMETRIC && embeddingDimension > 3)
            throw std::invalid_argument("the symmetric pair kernel is built for dimensions 1..3");
        kernel_ = kind;
    }

    void getTiming(double *out, size_t length) override {
        if (out == nullptr || length < 2) throw std::invalid_argument("timing output needs 2 values");
        out[0] = last_ms_[0];
        out[1] = last_ms_[1];
    }

  private:
    void invalidate() { ll_valid_ = grad_valid_ = false; }

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
        data_dirty_ = false;
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

    // Symmetric pass for D <= 3 once the data spans a few row blocks; the row-wise pass otherwise.
    bool use_symmetric() const {
        if (kernel_ == HPH_PAIR_KERNEL_ROWWISE) return false;
        if (embeddingDimension > 3) return false;
        if (kernel_ == HPH_PAIR_KERNEL_SYMMETRIC) return true;
        return locationCount >= kSymMinEvents;
    }

    // Runs the pair pass, finalize and reduction kernels for K parameter vectors; results land in
    // h_result_ ([K][7]).
    void run(const double *params, int K, bool grad, double *probs_host) {
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
        double *probs_dev = probs_host ? d_probs_.as<double>() : nullptr;
        if (use_symmetric()) {
            if (!sym_valid_ || sym_K_ != K) {
                sym_ = kern::plan_sym_grid(locationCount, embeddingDimension, K, kTargetBlocks);
                d_items_.ensure(sizeof(int2) * sym_.items.size());
                check(hipMemcpy(d_items_.ptr, sym_.items.data(), sizeof(int2) * sym_.items.size(),
                                hipMemcpyHostToDevice),
                      "upload work items");
                sym_valid_ = true;
                sym_K_ = K;
            }
            kern::SymGrid g = sym_.grid;
            g.item = d_items_.as<int2>();
            g.skip_zero = skip_zero;
            d_part_.ensure(sizeof(double) * kern::sym_rowpart_doubles(g, K, grad));
            d_colpart_.ensure(sizeof(double) * kern::sym_colpart_doubles(g, K, grad));
            if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            kern::launch_sym(ev, d_params_.as<double>(), K, g, grad, d_part_.as<double>(), d_colpart_.as<double>(),
                             stream_);
            check(hipGetLastError(), "symmetric pair kernel launch");
            if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
            kern::launch_sym_finalize(ev, d_params_.as<double>(), K, g, grad, d_part_.as<double>(),
                                      d_colpart_.as<double>(), d_block_.as<double>(), probs_dev, stream_);
            check(hipGetLastError(), "finalize kernel launch");
        } else {
            kern::PairGrid g = kern::plan_pair_grid(locationCount, K, kTargetBlocks);
            g.skip_zero = skip_zero;
            const int nacc = grad ? 6 : 2;
            d_part_.ensure(sizeof(double) * (size_t)K * g.chunks * nacc * npad_);
            if (timing_) check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            kern::launch_pair(ev, d_params_.as<double>(), K, g, grad, fp32_, d_part_.as<double>(), stream_);
            check(hipGetLastError(), "pair kernel launch");
            if (timing_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
            kern::launch_finalize(ev, d_params_.as<double>(), K, g, grad, d_part_.as<double>(),
                                  d_block_.as<double>(), probs_dev, stream_);
            check(hipGetLastError(), "finalize kernel launch");
        }
        kern::launch_reduce(ev, K, fin_blocks, grad, d_block_.as<double>(), d_result_.as<double>(), stream_);
        check(hipGetLastError(), "reduce kernel launch");
        if (timing_) check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
        check(hipMemcpyAsync(h_result_.ptr, d_result_.ptr, sizeof(double) * (size_t)K * kern::kResultWidth,
                             hipMemcpyDeviceToHost, stream_),
              "download result");
        if (probs_host)
            check(hipMemcpyAsync(probs_host, d_probs_.ptr, sizeof(double) * locationCount, hipMemcpyDeviceToHost,
                                 stream_),
                  "download probs");
        check(hipStreamSynchronize(stream_), "stream sync");
        if (timing_) {
            float a = 0.f, b = 0.f;
            check(hipEventElapsedTime(&a, ev_[0], ev_[1]), "hipEventElapsedTime");
            check(hipEventElapsedTime(&b, ev_[1], ev_[2]), "hipEventElapsedTime");
            last_ms_[0] = a;
            last_ms_[1] = b;
        }
    }

    void evaluate_single(bool grad, double *probs_host) {
        if (!have_params_) throw NotInitialized("parameters are not set (setParameters)");
        run(params_, 1, grad, probs_host);
        const double *r = h_result_.as<double>();
        ll_ = r[0];
        ll_valid_ = true;
        if (grad) {
            std::memcpy(grad_, r + 1, sizeof(grad_));
            grad_valid_ = true;
        }
    }

    static constexpr int kTargetBlocks = 4096;
    static constexpr int kSymMinEvents = 4096;
    int kernel_ = HPH_PAIR_KERNEL_AUTO;
    kern::SymPlan sym_;
    bool sym_valid_ = false;
    int sym_K_ = 0;

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
    bool ll_valid_ = false, grad_valid_ = false;

    bool has_stored_ = false, stored_have_params_ = false, stored_have_locs_ = false, stored_ll_valid_ = false;
    double stored_params_[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> stored_locs_;
    double stored_ll_ = 0.0;

    bool timing_ = false;
    hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};
    double last_ms_[2] = {0.0, 0.0};

    DevBuf d_x_, d_t_, d_perm_, d_probs_, d_params_, d_part_, d_block_, d_result_, d_colpart_, d_items_;
    HostBuf h_stage_, h_result_, h_params_;
};

SharedPtr factory(int dim1, int dim2, long flags, int device, int /*threads*/) {
    return std::make_shared<HipHawkes>(dim1, dim2, flags, device);
}

} // namespace hph

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

struct hph_engine {
    hph::SharedPtr impl;
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
        auto impl = hph::factory(embedding_dimension, location_count, flags, device, threads);
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

int hph_set_pair_kernel(hph_engine *e, int kind) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->setPairKernel(kind); });
}

int hph_get_timing(hph_engine *e, double *out, size_t length) {
    HPH_REQUIRE_ENGINE(e);
    return guarded([&] { e->impl->getTiming(out, length); });
}

} // extern "C"
