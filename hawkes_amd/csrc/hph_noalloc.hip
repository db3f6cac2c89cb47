// Allocation-free evaluation: hph_log_likelihood_noalloc (include/hph.h).
//
// For callers that cannot allocate device memory (code under HIP graph capture, serving loops):
// the log-likelihood and gradient of K parameter vectors of one dataset whose locations, times,
// parameters and result are already in device memory. Every intermediate -- the time-sorted
// structure-of-arrays copy of the events, the row-wise pass's partial sums, the finalize kernel's
// per-block double-double sums and a status word -- lives in `scratch`, one static device array of
// kScratchBytes per GPU (the runtime places it with the code object; it lives until the process
// ends). Nothing here calls hipMalloc, hipFree or an allocating library routine.
//
// The kernels are the row-wise pass of the engine (hph_kernels.hip: pair_kernel /
// pair_kernel_narrow, finalize_kernel, reduce_kernel) with the work decomposition the engine
// uses for a batch of K vectors with the row-wise pass at target grid 4096, so that the results
// equal hph_get_log_likelihood_batch's on such an engine bit for bit whenever that decomposition
// fits the array (otherwise fewer column chunks are used: the same sums in another grouping).
//
// Inputs are checked on the device (a kernel that cannot fault: every index is bounded by N and
// K): non-finite values or times that are not ascending set the status word, and a last kernel
// then writes NaN to every result.
#include "hph.h"
#include "hph_kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace hph {
namespace noalloc {

constexpr size_t kScratchBytes = (size_t)HPH_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
constexpr int kTarget = 4096; // target grid of the engine's row-wise pass (hph_set_pair_kernel default)

// The scratch array. Host code never uses this name as an address: kernels receive the pointer
// that hipGetSymbolAddress returns for the current device (scratch_on_device).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device, resolved once per device.
static unsigned char *scratch_on_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

__global__ void clear_status_kernel(int *status) { *status = 0; }

// SoA copy of the caller's row-major N x D locations and of the times, padded to npad with the
// last event (as the engine stages them), and the input checks: finite locations, times and
// parameters, times ascending. Thread s < max(npad, 6 K).
__global__ __launch_bounds__(256) void stage_kernel(int D, int n, int npad, const double *__restrict__ loc,
                                                    const double *__restrict__ times,
                                                    const double *__restrict__ params, int nparams,
                                                    double *__restrict__ x, double *__restrict__ t, int *status) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (s < npad) {
        const int src = s < n ? s : n - 1;
        for (int d = 0; d < D; ++d) {
            const double v = loc[(size_t)src * D + d];
            x[(size_t)d * npad + s] = v;
            bad |= !isfinite(v);
        }
        const double ts = times[src];
        t[s] = ts;
        bad |= !isfinite(ts);
        if (s + 1 < n) bad |= !(ts <= times[s + 1]);
    }
    if (s < nparams) bad |= !isfinite(params[s]);
    if (bad) atomicOr(status, 1);
}

// NaN into every result of the K vectors when the checks failed.
__global__ void poison_kernel(const int *status, int K, int width, double *result) {
    if (*status == 0) return;
    for (int q = threadIdx.x; q < K * width; q += blockDim.x) {
        const int k = q / width, c = q % width;
        result[(size_t)k * kern::kResultWidth + c] = __builtin_nan("");
    }
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Byte layout of the scratch array for one call.
struct Layout {
    size_t status = 0, x = 0, t = 0, part = 0, block = 0, end = 0;
};

static Layout layout_for(int D, int npad, int K, int nacc, int chunks, int row_blocks) {
    Layout L;
    L.status = 0;
    L.x = 256;
    L.t = L.x + align256(sizeof(double) * (size_t)D * npad);
    L.part = L.t + align256(sizeof(double) * (size_t)npad);
    L.block = L.part + align256(sizeof(double) * (size_t)K * chunks * nacc * npad);
    L.end = L.block + align256(sizeof(double) * (size_t)K * row_blocks * kern::kResultWidth * 2);
    return L;
}

void evaluate(int D, int N, const double *locations, const double *times, const double *params, int K, long flags,
              double *result, bool grad, hipStream_t stream) {
    if (D < 1 || D > kern::kMaxDim) throw std::invalid_argument("embedding_dimension must be in 1..8");
    if (N < 1) throw std::invalid_argument("location_count must be >= 1");
    if (N > (1 << 30)) throw std::invalid_argument("location_count must be at most 2^30");
    if (K < 1 || K > 65535) throw std::invalid_argument("K must be in 1..65535");
    if (locations == nullptr || times == nullptr || params == nullptr || result == nullptr)
        throw std::invalid_argument("a device pointer is null");
    const bool fp32 = (flags & HPH_FLAG_FLOAT) != 0;
    const int skip_zero = (flags & HPH_FLAG_ALL_PAIRS) ? 0 : 1;
    const int npad = (N + kern::kRowsPerBlock - 1) / kern::kRowsPerBlock * kern::kRowsPerBlock;
    const int nacc = grad ? 6 : 2;
    kern::PairGrid g = kern::plan_pair_grid(N, K, kTarget);
    Layout L = layout_for(D, npad, K, nacc, g.chunks, g.row_blocks);
    // fewer column chunks (a smaller target grid) until the partial sums fit
    for (int target = kTarget / 2; L.end > kScratchBytes && !g.narrow && g.chunks > 1 && target >= 1; target /= 2) {
        g = kern::plan_pair_grid(N, K, target);
        L = layout_for(D, npad, K, nacc, g.chunks, g.row_blocks);
    }
    if (L.end > kScratchBytes)
        throw std::invalid_argument("the evaluation needs " + std::to_string(L.end) + " bytes of scratch, more than the " +
                                    std::to_string(kScratchBytes) + " of the allocation-free entry point (N x K too large)");
    g.skip_zero = skip_zero;
    g.zero_arg_b = kern::background_skip_arg(N, fp32);
    g.zero_arg_s = kern::kZeroExpArg;
    g.cut_s = skip_zero;

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    unsigned char *base = scratch_on_device(dev);
    int *status = reinterpret_cast<int *>(base + L.status);
    double *x = reinterpret_cast<double *>(base + L.x);
    double *t = reinterpret_cast<double *>(base + L.t);
    double *part = reinterpret_cast<double *>(base + L.part);
    double *block = reinterpret_cast<double *>(base + L.block);

    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(1), 0, stream, status);
    const int nthreads = npad > 6 * K ? npad : 6 * K;
    hipLaunchKernelGGL(stage_kernel, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, stream, D, N, npad,
                       locations, times, params, 6 * K, x, t, status);
    kern::Events ev;
    ev.x = x;
    ev.t = t;
    ev.perm = nullptr; // only read where probabilities are written, and none are here
    ev.n = N;
    ev.npad = npad;
    ev.dim = D;
    ev.t_last = 0.0;
    ev.t_last_at = t + (N - 1); // T = the last (largest) time
    kern::launch_pair(ev, params, K, g, grad, fp32, part, stream);
    kern::launch_finalize(ev, params, K, g, grad, part, block, nullptr, nullptr, nullptr, stream);
    kern::launch_reduce(ev, K, g.row_blocks, grad, block, result, stream);
    hipLaunchKernelGGL(poison_kernel, dim3(1), dim3(256), 0, stream, status, K, grad ? kern::kResultWidth : 1, result);
    e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("allocation-free evaluation launch: ") + hipGetErrorString(e));
}

} // namespace noalloc
} // namespace hph
