/*
 * hph_engine.hpp -- C++ engine interface of the MI355X Hawkes likelihood engine.
 *
 * hph::AbstractHawkes restates src/AbstractHawkes.hpp:32-60: the same constructor, the same nine
 * pure virtuals and the same protected members, so that C++ callers written against the
 * reference recompile against this header, and so do engine subclasses written against the
 * reference's base class. The one difference: observationCount (N (N - 1) / 2, unused by the
 * reference's engines) is 64-bit here, because the reference's int overflows from N ~ 46,341 on.
 *
 * hph::HawkesEngine derives from it and adds this library's extensions (gradient, batched
 * proposals, timing, pass selection, autotuner, re-evaluation kind, location gradient).
 * hph::factory (src/AbstractHawkes.hpp:64, src/factory.cpp:32) returns the engine as the base
 * class, as the reference does; hph::engineFactory returns the same engine typed for the
 * extensions. Errors are reported as exceptions derived from std::exception
 * (std::invalid_argument for bad lengths or values, std::runtime_error for device failures);
 * the C ABI in hph.h turns them into status codes.
 */
#ifndef HPH_ENGINE_HPP
#define HPH_ENGINE_HPP

#include <cstddef>
#include <memory>
#include <vector>

namespace hph {

/* Flag bits (src/flags.h:6-13). */
enum Flags { FLOAT = 1 << 2, TBB = 1 << 3, OPENCL = 1 << 4, SSE = 1 << 5, AVX = 1 << 6, AVX512 = 1 << 7,
             ALL_PAIRS = 1 << 16 /* extension: evaluate every pair (no tile skipping) */ };

/* src/AbstractHawkes.hpp:32-60 */
class AbstractHawkes {
  public:
    AbstractHawkes(int embeddingDimension, int locationCount, long flags)
        : embeddingDimension(embeddingDimension), locationCount(locationCount),
          observationCount((long long)locationCount * (locationCount - 1) / 2), flags(flags) {}

    virtual ~AbstractHawkes() = default;

    // Interface
    virtual double getSumOfLikContribs() = 0;
    virtual void updateLocations(int, double *, size_t) = 0;
    virtual void getProbsSelfExcite(double *, size_t) = 0;
    virtual void storeState() = 0;
    virtual void restoreState() = 0;
    virtual void acceptState() = 0;
    virtual void setTimesData(double *, size_t) = 0;
    virtual void setParameters(double *, size_t) = 0;
    virtual int getInternalDimension() = 0;

  protected:
    int embeddingDimension;
    int locationCount;
    long long observationCount;
    long flags;

    int updatedLocation = -1;
};

typedef std::shared_ptr<hph::AbstractHawkes> SharedPtr;

/* The engine of this library: the reference interface plus extensions. */
class HawkesEngine : public AbstractHawkes {
  public:
    using AbstractHawkes::AbstractHawkes;

    // d loglik / d(sigmaXprec, tauXprec, tauTprec, omega, theta, mu0); the reference's
    // getLogLikelihoodGradient is commented out (src/AbstractHawkes.hpp:44).
    virtual void getLogLikeGradient(double *, size_t) = 0;
    virtual double getLogLikelihoodAndGradient(double *gradient6) = 0;
    // K parameter vectors (K x 6, row-major) evaluated in one launch.
    virtual void setParametersBatch(const double *, size_t) = 0;
    virtual void getLogLikelihoodBatch(double *loglik, double *gradient, size_t K) = 0;
    virtual void setTiming(bool enabled) = 0;
    virtual void setPairKernel(int kind) = 0;
    virtual void setPairConfig(int kind, int rows_per_lane, int target_blocks) = 0;
    virtual void autotune(int repeats, double *best_ms) = 0;
    virtual void getTunedConfig(int *out, size_t length) = 0;
    virtual void getTiming(double *out, size_t length) = 0;
    // How the last single-vector call was formed: 0 full pass, 1 excitation pass on cached
    // background sums, 2 cached sums only (HPH_EVAL_* in hph.h).
    virtual int getLastEvaluation() = 0;
    // Pair pass of the last full evaluation: out[0] = HPH_PAIR_KERNEL_ROWWISE, _SYMMETRIC or
    // _SPLIT, out[1] = 1 when it ran the base-2 form (scaled coordinates and times), else 0;
    // with length >= 3, out[2] = parameter vectors per launch; with length >= 4, out[3] = 1 when
    // the time-sorted symmetric pass ran the table exponential with the power of two in the table
    // entry (every background exponent of the call at most 700; bitwise the same results).
    virtual void getLastPass(int *out, size_t length) = 0;
    // Per-event sums (sum B, B r^2, B dt^2, S, S r^2, S dt; natural units, N x 6 row-major in the
    // caller's order) of the last evaluation with the gradient at the current parameters.
    virtual void getEventSums(double *out, size_t length) = 0;
    // d ll / d locations, N x D row-major in the caller's order; the reference declares a location
    // gradient for Java callers (src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96) without an
    // implementation.
    virtual void getLocationGradient(double *out, size_t length) = 0;
};

typedef std::shared_ptr<hph::HawkesEngine> EngineSharedPtr;

/* src/factory.cpp:32. Every flag combination returns the HIP engine (FLOAT selects the
 * FP32-bulk mode); device is the 0-based HIP device, -1 meaning device 0. */
SharedPtr factory(int dim1, int dim2, long flags, int device, int threads);

/* The same engine as factory(), typed for the extensions of HawkesEngine. */
EngineSharedPtr engineFactory(int dim1, int dim2, long flags, int device, int threads);

} // namespace hph

#endif // HPH_ENGINE_HPP
