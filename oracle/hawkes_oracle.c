/*
 * CPU oracle and CPU baselines for the spatio-temporal Hawkes log-likelihood.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this library, and only as the checker or as the timed
 * CPU baseline. The product (hawkes_amd, libhph.so) never links or calls it.
 *
 * Part 1 -- the oracle (oracle_impl.h, instantiated in long double and double):
 *   a restatement of src/NewHawkes.hpp:520-547, 636-660, 662-719 with compensated sums and
 *   the exact 6-parameter gradient.
 *   Pinned by: the JSON fixtures in tests/golden (50-digit mpmath evaluation of the independent R
 *   statement R/log_lik_and_grad.R:2-93 at tiny N), closed forms (N = 1, N = 2, theta = 0),
 *   invariants (translation, permutation, scaling), central finite differences and the
 *   R gradient formulas R/log_lik_and_grad.R:176-247 (mu0, tauX, tauT via the chain rule).
 *
 * Part 2 -- CPU baselines: restatements of the reference's three CPU flavours
 *   serial scalar  NewHawkes<DoubleNoSimdTypeInfo, CpuAccumulate>   (src/NewHawkes.hpp:899-904)
 *   AVX 4-wide     NewHawkes<DoubleAvxTypeInfo, CpuAccumulate>      (src/NewHawkes.hpp:927-931)
 *   threaded       NewHawkes<..., TbbAccumulate> (src/NewHawkes.hpp:812-847), OpenMP rows
 *   The per-pair arithmetic follows ratesLoop (src/NewHawkes.hpp:532-544): sqrt distance,
 *   three pdf() and one exp(-omega*dt) per pair, S computed for every j and masked.
 *   The +gradient variants add the six moment sums of the (corrected) seven-sum loop
 *   src/NewHawkes.hpp:592-634.
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- Part 1: oracle ---------------- */

#define REAL long double
#define SUFFIX _ld
#define EXP expl
#define LOG logl
#define SQRT sqrtl
#define ERFC erfcl
#define POW powl
#define FABS fabsl
#define INV_SQRT_2PI 0.398942280401432677939946059934381868L
#define INV_SQRT_2 0.707106781186547524400844362104849039L
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef EXP
#undef LOG
#undef SQRT
#undef ERFC
#undef POW
#undef FABS
#undef INV_SQRT_2PI
#undef INV_SQRT_2

#define REAL double
#define SUFFIX _d
#define EXP exp
#define LOG log
#define SQRT sqrt
#define ERFC erfc
#define POW pow
#define FABS fabs
#define INV_SQRT_2PI 0.398942280401432677939946059934381868
#define INV_SQRT_2 0.707106781186547524400844362104849039
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef EXP
#undef LOG
#undef SQRT
#undef ERFC
#undef POW
#undef FABS
#undef INV_SQRT_2PI
#undef INV_SQRT_2

/* precision: 0 = long double (default truth), 1 = double. Returns 0 on success. */
int hawkes_oracle_rows(int precision, int D, long N, const double *locs, const double *times,
                       const double *params, long row_begin, long row_end, double *out15) {
    if (D < 1 || N < 1 || row_begin < 0 || row_end > N || row_begin > row_end) return -1;
    if (precision == 0)
        oracle_rows_ld(D, N, locs, times, params, row_begin, row_end, out15, NULL);
    else
        oracle_rows_d(D, N, locs, times, params, row_begin, row_end, out15, NULL);
    return 0;
}

int hawkes_oracle_total(int precision, int D, long N, const double *locs, const double *times,
                        const double *params, double *ll, double *grad6, double *scale7, double *probs) {
    if (D < 1 || N < 1) return -1;
    if (precision == 0) return oracle_total_ld(D, N, locs, times, params, ll, grad6, scale7, probs);
    return oracle_total_d(D, N, locs, times, params, ll, grad6, scale7, probs);
}

/* Location gradient d ll / d x (N x D, caller's order) and its magnitude scale (may be NULL). */
int hawkes_oracle_location_gradient(int precision, int D, long N, const double *locs, const double *times,
                                    const double *params, double *grad, double *scale) {
    if (D < 1 || D > 8 || N < 1) return -1;
    if (precision == 0) return oracle_locgrad_ld(D, N, locs, times, params, grad, scale);
    return oracle_locgrad_d(D, N, locs, times, params, grad, scale);
}

int hawkes_oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------- Part 2: CPU baselines ---------------- */

#define M_1_SQRT_2PI_D 0.398942280401432677939946059934381868

static inline double bl_pdf(double z) { return M_1_SQRT_2PI_D * exp(-0.5 * z * z); }

/* Row terms shared by all flavours (src/NewHawkes.hpp:652-655). */
static inline double row_tail(double sum, double ti, double T, double omega, double theta, double mu0,
                              double tauT) {
    return log(sum) + theta * (exp(-omega * (T - ti)) - 1.0) -
           mu0 * (0.5 * erfc(-tauT * (T - ti) * M_SQRT1_2) - 0.5 * erfc(tauT * ti * M_SQRT1_2));
}

typedef struct {
    double sB, sBr, sBd, sS, sSr, sSd;
} moments_t;

/* Scalar pair loop (serial flavour). grad != 0 adds the moment sums. */
static double serial_row(int D, long N, const double *locs, const double *times, const double *p, long i,
                         int grad, moments_t *m) {
    const double sigma = p[0], tauX = p[1], tauT = p[2], omega = p[3], theta = p[4], mu0 = p[5];
    const double cB = mu0 * pow(tauX, D) * tauT, cS = pow(sigma, D) * theta * omega;
    const double ti = times[i];
    double sum = 0;
    moments_t acc = {0, 0, 0, 0, 0, 0};
    for (long j = 0; j < N; ++j) {
        double r2 = 0;
        for (int d = 0; d < D; ++d) {
            const double dx = locs[i * D + d] - locs[j * D + d];
            r2 += dx * dx;
        }
        const double r = sqrt(r2);
        const double dt = ti - times[j];
        const double b = bl_pdf(r * tauX) * bl_pdf(dt * tauT);
        const double s0 = exp(-omega * dt) * bl_pdf(r * sigma);
        const double s = dt > 0 ? s0 : 0.0;
        sum += cB * b + cS * s;
        if (grad) {
            acc.sB += b;
            acc.sBr += b * r2;
            acc.sBd += b * dt * dt;
            acc.sS += s;
            acc.sSr += s * r2;
            acc.sSd += s * dt;
        }
    }
    if (m) *m = acc;
    return sum;
}

/* libmvec 4-wide FP64 exp (AVX2 variant), as the AVX engine uses a vector exp. */
__m256d _ZGVdN4v_exp(__m256d x);

__attribute__((target("avx2,fma"))) static double avx_row(int D, long N, const double *locs,
                                                          const double *times, const double *p, long i,
                                                          int grad, moments_t *m) {
    const double sigma = p[0], tauX = p[1], tauT = p[2], omega = p[3], theta = p[4], mu0 = p[5];
    const double cB = mu0 * pow(tauX, D) * tauT, cS = pow(sigma, D) * theta * omega;
    const __m256d vcB = _mm256_set1_pd(cB), vcS = _mm256_set1_pd(cS);
    const __m256d vti = _mm256_set1_pd(times[i]);
    const __m256d half = _mm256_set1_pd(-0.5), k = _mm256_set1_pd(M_1_SQRT_2PI_D);
    const __m256d vtx = _mm256_set1_pd(tauX), vtt = _mm256_set1_pd(tauT), vsg = _mm256_set1_pd(sigma);
    const __m256d vom = _mm256_set1_pd(-omega), zero = _mm256_setzero_pd();
    __m256d vsum = zero, a0 = zero, a1 = zero, a2 = zero, a3 = zero, a4 = zero, a5 = zero;
    const long vec = N - N % 4;
    for (long j = 0; j < vec; j += 4) {
        /* gathered distance, as DistanceDispatch builds SIMD lanes from row-major locations */
        __m256d r2 = zero;
        for (int d = 0; d < D; ++d) {
            const __m256d xj = _mm256_set_pd(locs[(j + 3) * D + d], locs[(j + 2) * D + d],
                                             locs[(j + 1) * D + d], locs[j * D + d]);
            const __m256d dx = _mm256_sub_pd(_mm256_set1_pd(locs[i * D + d]), xj);
            r2 = _mm256_fmadd_pd(dx, dx, r2);
        }
        const __m256d r = _mm256_sqrt_pd(r2);
        const __m256d dt = _mm256_sub_pd(vti, _mm256_loadu_pd(times + j));
        const __m256d zx = _mm256_mul_pd(r, vtx), zt = _mm256_mul_pd(dt, vtt), zs = _mm256_mul_pd(r, vsg);
        const __m256d px = _mm256_mul_pd(k, _ZGVdN4v_exp(_mm256_mul_pd(half, _mm256_mul_pd(zx, zx))));
        const __m256d pt = _mm256_mul_pd(k, _ZGVdN4v_exp(_mm256_mul_pd(half, _mm256_mul_pd(zt, zt))));
        const __m256d ps = _mm256_mul_pd(k, _ZGVdN4v_exp(_mm256_mul_pd(half, _mm256_mul_pd(zs, zs))));
        const __m256d eo = _ZGVdN4v_exp(_mm256_mul_pd(vom, dt));
        const __m256d b = _mm256_mul_pd(px, pt);
        const __m256d s = _mm256_and_pd(_mm256_cmp_pd(dt, zero, _CMP_GT_OQ), _mm256_mul_pd(eo, ps));
        vsum = _mm256_add_pd(vsum, _mm256_fmadd_pd(vcB, b, _mm256_mul_pd(vcS, s)));
        if (grad) {
            a0 = _mm256_add_pd(a0, b);
            a1 = _mm256_fmadd_pd(b, r2, a1);
            a2 = _mm256_fmadd_pd(b, _mm256_mul_pd(dt, dt), a2);
            a3 = _mm256_add_pd(a3, s);
            a4 = _mm256_fmadd_pd(s, r2, a4);
            a5 = _mm256_fmadd_pd(s, dt, a5);
        }
    }
    double t[4];
    double sum;
    _mm256_storeu_pd(t, vsum);
    sum = (t[0] + t[1]) + (t[2] + t[3]);
    moments_t acc = {0, 0, 0, 0, 0, 0};
    if (grad) {
        double *f[6] = {&acc.sB, &acc.sBr, &acc.sBd, &acc.sS, &acc.sSr, &acc.sSd};
        __m256d v[6] = {a0, a1, a2, a3, a4, a5};
        for (int q = 0; q < 6; ++q) {
            _mm256_storeu_pd(t, v[q]);
            *f[q] = (t[0] + t[1]) + (t[2] + t[3]);
        }
    }
    /* scalar tail, as the reference's edge-case loop (src/NewHawkes.hpp:645-650) */
    for (long j = vec; j < N; ++j) {
        double r2 = 0;
        for (int d = 0; d < D; ++d) {
            const double dx = locs[i * D + d] - locs[j * D + d];
            r2 += dx * dx;
        }
        const double r = sqrt(r2);
        const double dt = times[i] - times[j];
        const double b = bl_pdf(r * tauX) * bl_pdf(dt * tauT);
        const double s = dt > 0 ? exp(-omega * dt) * bl_pdf(r * sigma) : 0.0;
        sum += cB * b + cS * s;
        if (grad) {
            acc.sB += b;
            acc.sBr += b * r2;
            acc.sBd += b * dt * dt;
            acc.sS += s;
            acc.sSr += s * r2;
            acc.sSd += s * dt;
        }
    }
    if (m) *m = acc;
    return sum;
}

/* Converts one row's sums into ll_i and its gradient terms. */
static void finish_row(int D, long N, const double *times, const double *p, long i, double sum,
                       const moments_t *m, double *out7) {
    const double sigma = p[0], tauX = p[1], tauT = p[2], omega = p[3], theta = p[4], mu0 = p[5];
    const double T = times[N - 1], ti = times[i], dT = T - ti;
    out7[0] = row_tail(sum, ti, T, omega, theta, mu0, tauT);
    if (!m) return;
    const double sigmaD = pow(sigma, D), tauXD = pow(tauX, D);
    const double eT = exp(-omega * dT);
    const double dl[6] = {
        theta * omega * pow(sigma, D - 1) * (D * m->sS - sigma * sigma * m->sSr),
        mu0 * tauT * pow(tauX, D - 1) * (D * m->sB - tauX * tauX * m->sBr),
        mu0 * tauXD * (m->sB - tauT * tauT * m->sBd),
        theta * sigmaD * (m->sS - omega * m->sSd),
        omega * sigmaD * m->sS,
        tauXD * tauT * m->sB,
    };
    const double di[6] = {
        0,
        0,
        -mu0 * (dT * bl_pdf(tauT * dT) + ti * bl_pdf(tauT * ti)),
        -theta * dT * eT,
        eT - 1,
        -(0.5 * erfc(-tauT * dT * M_SQRT1_2) - 0.5 * erfc(tauT * ti * M_SQRT1_2)),
    };
    for (int k = 0; k < 6; ++k) out7[1 + k] = dl[k] / sum + di[k];
}

/*
 * flavour: 0 = serial scalar, 1 = AVX2 4-wide, 2 = threaded (OpenMP rows) + AVX2.
 * Evaluates rows [row_begin, row_end) against all N events (a bounded sample of a full call
 * when the range is a subset). out7[0] = sum of ll_i over the rows; out7[1..6] = gradient
 * partials when grad != 0. The N*(D-1)*log(1/sqrt(2 pi)) constant is added when the range
 * covers every row. Returns 0 on success.
 */
int hawkes_cpu_baseline(int flavour, int grad, int D, long N, const double *locs, const double *times,
                        const double *params, long row_begin, long row_end, double *out7) {
    if (D < 1 || N < 1 || row_begin < 0 || row_end > N || row_begin > row_end) return -1;
    double tot[7] = {0, 0, 0, 0, 0, 0, 0};
    if (flavour == 2) {
        const long n = row_end - row_begin;
        double *rows = (double *)malloc(sizeof(double) * 7 * (size_t)(n > 0 ? n : 1));
        if (!rows) return -1;
#pragma omp parallel for schedule(dynamic, 8)
        for (long i = row_begin; i < row_end; ++i) {
            moments_t m;
            const double s = avx_row(D, N, locs, times, params, i, grad, grad ? &m : NULL);
            finish_row(D, N, times, params, i, s, grad ? &m : NULL, rows + (i - row_begin) * 7);
        }
        for (long q = 0; q < n; ++q)
            for (int k = 0; k < (grad ? 7 : 1); ++k) tot[k] += rows[q * 7 + k];
        free(rows);
    } else {
        for (long i = row_begin; i < row_end; ++i) {
            moments_t m;
            double r7[7];
            const double s = flavour == 1 ? avx_row(D, N, locs, times, params, i, grad, grad ? &m : NULL)
                                          : serial_row(D, N, locs, times, params, i, grad, grad ? &m : NULL);
            finish_row(D, N, times, params, i, s, grad ? &m : NULL, r7);
            for (int k = 0; k < (grad ? 7 : 1); ++k) tot[k] += r7[k];
        }
    }
    if (row_begin == 0 && row_end == N) tot[0] += (double)N * (D - 1) * log(M_1_SQRT_2PI_D);
    memcpy(out7, tot, sizeof(double) * (grad ? 7 : 1));
    return 0;
}
