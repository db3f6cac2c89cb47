/*
 * hph.h -- C ABI of the MI355X Hawkes likelihood engine (libhph.so).
 *
 * Drop-in boundary for the engine API of suchard-group/hawkes (hpHawkes):
 * every entry point below replaces one virtual of hph::AbstractHawkes
 * (src/AbstractHawkes.hpp:32-60) or the factory (src/factory.cpp:32-118), as a
 * foreign binding (Rcpp .Call, JNI, ctypes, cgo) would bind it. Plain pointers and
 * sizes only; the caller owns every buffer, which is borrowed for the call only.
 *
 * Every function returns HPH_OK (0) or an error code; hph_last_error() gives the
 * message of the calling thread's last failure. No call aborts the process.
 * One engine must not be used by two threads at once; distinct engines are independent.
 */
#ifndef HPH_H
#define HPH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flag bits, same values as hph::Flags (src/flags.h:6-13). */
#define HPH_FLAG_FLOAT (1L << 2)  /* FP32-bulk pair arithmetic, FP64 accumulation */
#define HPH_FLAG_TBB (1L << 3)    /* accepted; CPU-only, no effect here */
#define HPH_FLAG_OPENCL (1L << 4) /* GPU engine: this library's HIP engine */
#define HPH_FLAG_SSE (1L << 5)    /* accepted; CPU-only, no effect here */
#define HPH_FLAG_AVX (1L << 6)    /* accepted; CPU-only, no effect here */
#define HPH_FLAG_AVX512 (1L << 7) /* accepted; CPU-only, no effect here */
/* Extension: evaluate every pair. By default a likelihood or gradient pass skips tiles of
 * time-sorted events whose background terms are all below e^-(45 + ln N) and whose excitation
 * terms are all below e^-X_S, X_S = 50 + ln N + max(0, ln(|cS| / |cB|)) (cB = mu0 tauX^D tauT /
 * 2 pi, cS = theta omega sigmaX^D / sqrt(2 pi)); the skipped terms move every event's background
 * sums by under 4e-18 and its rate by under 2e-22 relative, so results agree with this flag's to
 * FP64 rounding, not bitwise.
 * The excitation cut is made only where cB and cS have the same sign. A probability request
 * (their likelihood cut is bounded relative to the rate, not to a tiny probability) skips exactly
 * zero excitation terms only, or, where the excitation sums come from the band pass (the split
 * passes, and after a likelihood call at the same tauXprec and tauTprec), stops each event's sum
 * once the terms left could change it by at most 2^-60 of itself. */
#define HPH_FLAG_ALL_PAIRS (1L << 16)

/* Status codes. */
#define HPH_OK 0
#define HPH_ERR_INVALID_ARGUMENT 1
#define HPH_ERR_NOT_INITIALIZED 2
#define HPH_ERR_DEVICE 3
#define HPH_ERR_OUT_OF_MEMORY 4
#define HPH_ERR_UNSUPPORTED 5

/* Number of model parameters: (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0). */
#define HPH_PARAMETER_COUNT 6

typedef struct hph_engine hph_engine;

/* Message of the calling thread's last failed call ("" if none). */
const char *hph_last_error(void);

/* Replaces hph::factory(dim, N, flags, device, threads) (src/factory.cpp:32) and the engine
 * constructor (src/AbstractHawkes.hpp:34). device is the 0-based HIP device; -1 means device 0
 * (the reference's "no device" value, src/hpHawkes.cpp:65). threads is accepted and unused. */
int hph_create(int embedding_dimension, int location_count, long flags, int device, int threads,
               hph_engine **out);

/* Frees the engine and its device memory (replaces the shared_ptr release). NULL is a no-op. */
int hph_destroy(hph_engine *engine);

/* Replaces AbstractHawkes::setTimesData(double*, size_t) (src/AbstractHawkes.hpp:49).
 * length must equal location_count. Times need not be sorted; T is the last element. */
int hph_set_times_data(hph_engine *engine, const double *data, size_t length);

/* Replaces AbstractHawkes::updateLocations(int, double*, size_t) (src/AbstractHawkes.hpp:43).
 * location_index == -1: all locations, row-major N x D, length == N*D.
 * location_index >= 0 : one location, length == D. */
int hph_update_locations(hph_engine *engine, int location_index, const double *location, size_t length);

/* Replaces AbstractHawkes::setParameters(double*, size_t) (src/AbstractHawkes.hpp:50).
 * length must be 6: (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0); values finite. */
int hph_set_parameters(hph_engine *engine, const double *data, size_t length);

/* Replaces AbstractHawkes::getSumOfLikContribs() (src/AbstractHawkes.hpp:42): the log-likelihood. */
int hph_get_sum_of_lik_contribs(hph_engine *engine, double *out);

/* New (the reference's getLogLikelihoodGradient is commented out, src/AbstractHawkes.hpp:44):
 * d loglik / d(sigmaXprec, tauXprec, tauTprec, omega, theta, mu0); length must be 6. */
int hph_get_log_like_gradient(hph_engine *engine, double *out, size_t length);

/* Log-likelihood and gradient from one pass: *loglik and grad[0..5]. */
int hph_get_log_likelihood_and_gradient(hph_engine *engine, double *loglik, double *grad);

/* Replaces AbstractHawkes::getProbsSelfExcite(double*, size_t) (src/AbstractHawkes.hpp:45);
 * length must equal location_count; result in the caller's event order. */
int hph_get_probs_self_excite(hph_engine *engine, double *out, size_t length);

/* d loglik / d locations: out is N x D row-major in the caller's event order, length = N * D.
 * No counterpart in the reference's engine: its Java binding declares getLocationGradient
 * (src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96) and leaves it unimplemented
 * (src/jni/dr_inference_hawkes_NativeHPHSingleton.cpp:80-89). FP64 in every precision mode.
 * With timing on, hph_get_timing reports the rate pass and the gradient pass. */
int hph_get_location_gradient(hph_engine *engine, double *out, size_t length);

/* Replace storeState / restoreState / acceptState (src/AbstractHawkes.hpp:46-48): snapshot of
 * parameters, locations and the last log-likelihood; restore brings the snapshot back. */
int hph_store_state(hph_engine *engine);
int hph_restore_state(hph_engine *engine);
int hph_accept_state(hph_engine *engine);

/* Replaces AbstractHawkes::getInternalDimension() (src/AbstractHawkes.hpp:51). */
int hph_get_internal_dimension(hph_engine *engine, int *out);

/* ---- Batched proposals (extension; no counterpart in the reference) ----
 * K parameter vectors evaluated against the engine's data in one launch. params is K x 6,
 * row-major; count must be K*6. The single-vector parameters set by hph_set_parameters are
 * not changed. */
int hph_set_parameters_batch(hph_engine *engine, const double *params, size_t count);
/* loglik[K]; grad (K x 6, may be NULL for log-likelihood only). When every vector has the same
 * tauXprec and tauTprec (the scan set of the reference's sampler, R/Untitled.R:254-422: sigmaXprec,
 * omega, theta, mu0), the background sums -- which depend on the data, tauXprec and tauTprec
 * alone -- are formed once (or taken from the single-vector cache below) and each vector runs an
 * excitation pass only; hph_get_last_evaluation then reports HPH_EVAL_EXCITATION, else
 * HPH_EVAL_FULL. Dimensions 1..3, K >= 2, not with HPH_FLAG_ALL_PAIRS. */
int hph_get_log_likelihood_batch(hph_engine *engine, double *loglik, double *grad, size_t K);

/* ---- Kernel timing (extension; the reference's MICRO_BENCHMARK timers,
 * src/backend/opencl/OpenCLHawkes.hpp:38, 377-429, measured here with HIP events on the
 * engine's stream) ----
 * enabled != 0 records events around the kernels of every later evaluation.
 * hph_get_timing: out[0] = pair-pass kernel ms of the last evaluation, out[1] = per-row
 * finalize + reduction kernels ms; length must be >= 2. */
int hph_set_timing(hph_engine *engine, int enabled);
int hph_get_timing(hph_engine *engine, double *out, size_t length);

/* ---- Re-evaluation from cached per-event sums (extension; no counterpart in the reference) ----
 * A single-vector log-likelihood (or gradient) call keeps the per-event background and
 * excitation sums of its parameters. A later call whose parameters differ only in sigmaXprec,
 * omega, theta and mu0 (the parameters the reference's random-scan sampler proposes,
 * R/Untitled.R:254-422) reuses the background sums and runs an excitation-only pass, or, when
 * only theta and mu0 changed, reuses both (O(N)). Data changes, tauXprec or tauTprec changes and
 * HPH_FLAG_ALL_PAIRS take the full pass; dimensions 1..3. A probability request at the cached
 * tauXprec and tauTprec reuses the background sums and forms the excitation sums anew, each to
 * 2^-60 of itself (it leaves the cache as it was). A batch whose vectors share tauXprec and
 * tauTprec reuses (or refills) the cached background sums too.
 * hph_get_last_evaluation: *kind = HPH_EVAL_FULL, HPH_EVAL_EXCITATION (cached background sums +
 * excitation pass) or HPH_EVAL_CACHED (cached sums only) for the last single-vector call or
 * batch. */
#define HPH_EVAL_FULL 0
#define HPH_EVAL_EXCITATION 1
#define HPH_EVAL_CACHED 2
int hph_get_last_evaluation(hph_engine *engine, int *kind);
/* Pair pass of the last full evaluation (a call that re-evaluated from cached sums ran none):
 * out[0] = HPH_PAIR_KERNEL_ROWWISE, HPH_PAIR_KERNEL_SYMMETRIC or HPH_PAIR_KERNEL_SPLIT; out[1] = 1
 * when it ran the base-2 form (coordinates and times scaled so that the background term is 2^-y
 * of a sum of squares; chosen only where the rounding of the scaled values stays below 1e-13
 * relative on the terms), 0 for the natural form; with length >= 3, out[2] = parameter vectors
 * per launch (a batch whose symmetric or split partial sums do not fit in device memory at once
 * runs as consecutive launches of that many vectors); with length >= 4, out[3] = 1 when the
 * time-sorted symmetric pass ran in FP64 with every background exponent of the call at most 700
 * (from the data's extents and each vector's tauX, tauT), where its table exponential puts the
 * power of two into the table entry instead of scaling the result (bitwise the same results).
 * length must be >= 2; HPH_ERR_NOT_INITIALIZED before the first evaluation. */
int hph_get_last_pass(hph_engine *engine, int *out, size_t length);
/* The per-event sums behind the last log-likelihood-and-gradient evaluation at the current
 * parameters (dimensions 1..3): out is N x 6 row-major in the caller's order, length = 6 N, with
 * sum_j B_ij, sum_j B_ij r_ij^2, sum_j B_ij dt_ij^2, sum_j S_ij, sum_j S_ij r_ij^2, sum_j S_ij dt_ij,
 * B_ij = exp(-(tauX^2 r_ij^2 + tauT^2 dt_ij^2) / 2) over all j (j = i included) and
 * S_ij = exp(-omega dt_ij - sigmaX^2 r_ij^2 / 2) over t_j < t_i (the rate is
 * lambda_i = mu0 tauX^D tauT sum B / 2 pi + theta omega sigmaX^D sum S / sqrt(2 pi)), within the
 * likelihood passes' negligible-term cuts. HPH_ERR_NOT_INITIALIZED when no such evaluation is
 * at hand. */
int hph_get_event_sums(hph_engine *engine, double *out, size_t length);

/* ---- Pair-pass selection (extension) ----
 * AUTO: the tuned or built-in choice among the others (row-wise below ~12k events or above
 * dimension 3; symmetric above; split for dimensions 1..3, at least 65,536 events and locations
 * spread over at least 3 times the distance at which background terms become negligible).
 * ROWWISE: every ordered pair, one event row per lane. SYMMETRIC: background terms once per
 * unordered pair (dimensions 1..3). SPLIT: the symmetric background terms in a pass of their own
 * over events reordered within time windows along a space-filling curve, skipping groups whose
 * bounding boxes are far apart, then the excitation terms in a second pass (dimensions 1..3).
 * Results agree to rounding; each is deterministic. */
#define HPH_PAIR_KERNEL_AUTO 0
#define HPH_PAIR_KERNEL_ROWWISE 1
#define HPH_PAIR_KERNEL_SYMMETRIC 2
#define HPH_PAIR_KERNEL_SPLIT 3
int hph_set_pair_kernel(hph_engine *engine, int kind);
/* Full form: rows_per_lane (1 or 2) applies to the symmetric pass; target_blocks (64..2^20) is
 * the grid size the work decomposition aims for. hph_set_pair_kernel(e, k) is
 * hph_set_pair_config(e, k, 2, 4096). */
int hph_set_pair_config(hph_engine *engine, int kind, int rows_per_lane, int target_blocks);

/* ---- Autotuner (extension) ----
 * Times every pair-pass configuration (row-wise; symmetric with 1 or 2 rows per lane; four
 * grid sizes) on the engine's current data and parameters (the batch when one is set), keeps
 * the fastest for this shape (log2 N, dimension, batch size, precision) for every engine of the
 * process, and appends it to the file named by the HPH_TUNE_CACHE environment variable, if set,
 * which later processes read. best_ms (may be NULL) receives the winning kernel time.
 * hph_get_tuned_config: out[0] = pair kernel kind, out[1] = rows per lane, out[2] = target grid
 * size of the configuration a single-vector evaluation would use now; length must be >= 3. */
int hph_autotune(hph_engine *engine, int repeats, double *best_ms);
int hph_get_tuned_config(hph_engine *engine, int *out, size_t length);

/* ---- Dataset batches (extension; no counterpart in the reference) ----
 * Many independent datasets of uneven size (dimension 1..3), each with its own parameter
 * vectors, evaluated by one persistent launch over a cost-ordered work queue. Flags as in
 * hph_create (FLOAT: FP32-bulk; ALL_PAIRS). */
typedef struct hph_batch hph_batch;
int hph_batch_create(int embedding_dimension, int dataset_count, const int *location_counts, long flags, int device,
                     hph_batch **out);
int hph_batch_destroy(hph_batch *batch);
/* Dataset d: locations row-major N_d x D (length N_d*D), times length N_d (T = last element). */
int hph_batch_set_data(hph_batch *batch, int dataset, const double *locations, size_t locations_length,
                       const double *times, size_t times_length);
/* vectors_per_dataset[d] >= 1; params: the vectors of dataset 0, then dataset 1, ..., each 6
 * values; count = 6 x total vectors. */
int hph_batch_set_parameters(hph_batch *batch, const int *vectors_per_dataset, const double *params, size_t count);
/* loglik[total_vectors]; grad (total_vectors x 6, may be NULL), in the order of the parameters. */
int hph_batch_get_log_likelihood(hph_batch *batch, double *loglik, double *grad, size_t total_vectors);
/* enabled != 0: capture the launch sequence in a HIP graph once per shape and replay it. */
int hph_batch_set_graph(hph_batch *batch, int enabled);
/* Work decomposition: rows per lane (1 or 2) and the target number of work items, fixed (overrides
 * the tuned and the built-in choice: 2 rows per lane, 8192 items). */
int hph_batch_set_config(hph_batch *batch, int rows_per_lane, int target_items);
/* Times the work decompositions (1 or 2 rows per lane x six target item counts) on the current
 * datasets and vectors, keeps the fastest for this table's shape (log2 of sum_d K_d N_d^2,
 * dimension, total vectors, precision, dataset count) in the process-wide tuning cache shared with
 * hph_autotune (and the HPH_TUNE_CACHE file), and clears any fixed configuration: later
 * evaluations of a table of this shape use it. best_ms (may be NULL): the winner's device time. */
int hph_batch_autotune(hph_batch *batch, int repeats, double *best_ms);
/* The decomposition the next evaluation uses: out[0] rows per lane, out[1] target items,
 * out[2] = 0 built-in, 1 tuned, 2 set by hph_batch_set_config; length >= 3. */
int hph_batch_get_config(hph_batch *batch, int *out, size_t length);
/* Device time (ms) of the last evaluation's launch sequence. */
int hph_batch_get_timing(hph_batch *batch, double *out);
/* *out = 1 when dataset d took the split passes in the last evaluation (data spread over several
 * negligible-background distances for each of its vectors, at least 65,536 events, as for an
 * engine's built-in choice; the other datasets share one time-sorted pass), else 0. */
int hph_batch_last_split(hph_batch *batch, int dataset, int *out);

/* ---- Allocation-free evaluation (extension) ----
 * No counterpart in the reference: for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The log-likelihood
 * (and gradient) of K parameter vectors of one dataset, as hph_get_log_likelihood_batch returns
 * them, from inputs and an output that are already in device memory (pointers valid on the
 * current HIP device):
 *   locations: N x D row-major (the reference's layout, src/NewHawkes.hpp:244-276);
 *   times: N values in ascending order (T = the last); params: K x 6, in the order of
 *   hph_set_parameters; result: K x 7 (ll, then d ll / d params when gradient != 0; without the
 *   gradient only the first value of each 7 is written).
 * flags: HPH_FLAG_FLOAT and HPH_FLAG_ALL_PAIRS as in hph_create. The work is enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) and the call returns without waiting. Every intermediate
 * lives in one static device array of HPH_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; the call allocates nothing (the first call on a device looks up that array's address;
 * make it outside any graph capture). Limit: the events' copy plus the row-wise pass's partial
 * sums, about 8 (D + 1 + 6 K c) N bytes with c >= 1 column chunks, must fit in the array (with the
 * gradient, K x N up to about 1.3e6); larger requests return HPH_ERR_INVALID_ARGUMENT.
 * Non-finite inputs or times out of order give NaN in every result (checked on the device).
 * All calls on one GPU share the array, which keeps the last call's contents: calls that overlap
 * in time on one GPU race -- serialise them (one stream, or events between streams). */
#define HPH_NOALLOC_SCRATCH_BYTES (64u << 20)
int hph_log_likelihood_noalloc(int embedding_dimension, int location_count, const double *locations,
                               const double *times, const double *params, int K, long flags, double *result,
                               int gradient, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HPH_H */
