/*
 * Oracle core, instantiated twice by hawkes_oracle.c: once with REAL = long double
 * (the truth used by the parity tests) and once with REAL = double.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker, never the product.
 *
 * Model (reference semantics, file:line in suchard-group/hawkes):
 *   rate_ij = mu0*tauX^D*tauT * pdf(tauX*r_ij)*pdf(tauT*dt_ij)
 *           + sigma^D*theta*omega * [dt_ij > 0] * exp(-omega*dt_ij) * pdf(sigma*r_ij)
 *       -- src/NewHawkes.hpp:520-547 (ratesLoop): all j in [0,N), j == i included,
 *          strict dt > 0 mask (:540).
 *   r_ij = sqrt(sum_d (x_id - x_jd)^2)        -- src/Distance.hpp:72-83
 *   pdf(z) = exp(-z^2/2)/sqrt(2 pi)           -- src/NewHawkes.hpp:108-111
 *   Phi(z) = 0.5*erfc(-z/sqrt 2)              -- src/CDF.h:583-592 (used as exp(phi_new))
 *   ll_i = log(sum_j rate_ij) + theta*(exp(-omega*(T-t_i)) - 1)
 *          - mu0*(Phi(tauT*(T-t_i)) - Phi(-tauT*t_i)),   T = t[N-1]
 *       -- src/NewHawkes.hpp:652-655
 *   ll = sum_i ll_i + N*(D-1)*log(1/sqrt(2 pi))          -- src/NewHawkes.hpp:659
 *   p_i = S_i/(B_i + S_i) with the constant-scaled sums  -- src/NewHawkes.hpp:685, 707-714
 *
 * The gradient has no live counterpart in the reference (the C++ one is commented out,
 * src/NewHawkes.hpp:341-402); it is the exact derivative of ll above.
 * Sums over j and over i use Neumaier compensation.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)

typedef struct {
    REAL s, c;
} FN(nsum_t);

static inline void FN(nsum_add)(FN(nsum_t) * a, REAL x) {
    REAL t = a->s + x;
    if (FABS(a->s) >= FABS(x))
        a->c += (a->s - t) + x;
    else
        a->c += (x - t) + a->s;
    a->s = t;
}

static inline REAL FN(nsum_get)(const FN(nsum_t) * a) { return a->s + a->c; }

static inline REAL FN(pdf)(REAL z) { return EXP(-z * z / 2) * INV_SQRT_2PI; }

static inline REAL FN(cdf)(REAL z) { return ERFC(-z * INV_SQRT_2) / 2; }

/*
 * Per-row statistics for rows [row_begin, row_end). out has 15 REAL-converted doubles per row:
 *   0..5  sB, sB*r^2, sB*dt^2, sS, sS*r^2, sS*dt   (pdf-form sums: b_ij = pdf(tauX r) pdf(tauT dt),
 *                                                  s_ij = [dt>0] exp(-omega dt) pdf(sigma r))
 *   6     lambda_i (constant-scaled rate)
 *   7     ll_i (log lambda_i + integral terms)
 *   8     p_i (self-excitation probability)
 *   9..14 d ll_i / d(sigma, tauX, tauT, omega, theta, mu0)
 */
static void FN(oracle_rows)(int D, long N, const double *locs, const double *times, const double *params,
                            long row_begin, long row_end, double *out, double *abs_out) {
    const REAL sigma = params[0], tauX = params[1], tauT = params[2];
    const REAL omega = params[3], theta = params[4], mu0 = params[5];
    const REAL T = times[N - 1];
    const REAL sigmaD = POW(sigma, (REAL)D), tauXD = POW(tauX, (REAL)D);
    const REAL cB = mu0 * tauXD * tauT;
    const REAL cS = sigmaD * theta * omega;

#pragma omp parallel for schedule(dynamic, 4)
    for (long i = row_begin; i < row_end; ++i) {
        FN(nsum_t) acc[6] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
        const REAL ti = times[i];
        for (long j = 0; j < N; ++j) {
            REAL r2 = 0;
            for (int d = 0; d < D; ++d) {
                const REAL dx = (REAL)locs[i * D + d] - (REAL)locs[j * D + d];
                r2 += dx * dx;
            }
            const REAL r = SQRT(r2);
            const REAL dt = ti - (REAL)times[j];
            const REAL b = FN(pdf)(tauX * r) * FN(pdf)(tauT * dt);
            FN(nsum_add)(&acc[0], b);
            FN(nsum_add)(&acc[1], b * r2);
            FN(nsum_add)(&acc[2], b * dt * dt);
            if (dt > 0) {
                const REAL s = EXP(-omega * dt) * FN(pdf)(sigma * r);
                FN(nsum_add)(&acc[3], s);
                FN(nsum_add)(&acc[4], s * r2);
                FN(nsum_add)(&acc[5], s * dt);
            }
        }
        REAL sB = FN(nsum_get)(&acc[0]), sBr = FN(nsum_get)(&acc[1]), sBd = FN(nsum_get)(&acc[2]);
        REAL sS = FN(nsum_get)(&acc[3]), sSr = FN(nsum_get)(&acc[4]), sSd = FN(nsum_get)(&acc[5]);
        const REAL bg = cB * sB, se = cS * sS;
        const REAL lam = bg + se;
        const REAL dT = T - ti;
        const REAL eT = EXP(-omega * dT);
        const REAL phiA = FN(cdf)(tauT * dT), phiB = FN(cdf)(-tauT * ti);
        const REAL ll = LOG(lam) + theta * (eT - 1) - mu0 * (phiA - phiB);

        /* d lambda / d p (exact derivative of the rate above) */
        const REAL dl[6] = {
            theta * omega * POW(sigma, (REAL)(D - 1)) * (D * sS - sigma * sigma * sSr),
            mu0 * tauT * POW(tauX, (REAL)(D - 1)) * (D * sB - tauX * tauX * sBr),
            mu0 * tauXD * (sB - tauT * tauT * sBd),
            theta * sigmaD * (sS - omega * sSd),
            omega * sigmaD * sS,
            tauXD * tauT * sB,
        };
        /* d(integral terms)/d p */
        const REAL di[6] = {
            0,
            0,
            -mu0 * (dT * FN(pdf)(tauT * dT) + ti * FN(pdf)(tauT * ti)),
            -theta * dT * eT,
            eT - 1,
            -(phiA - phiB),
        };
        double *o = out + (i - row_begin) * 15;
        o[0] = (double)sB; o[1] = (double)sBr; o[2] = (double)sBd;
        o[3] = (double)sS; o[4] = (double)sSr; o[5] = (double)sSd;
        o[6] = (double)lam;
        o[7] = (double)ll;
        o[8] = (double)(se / (bg + se));
        for (int k = 0; k < 6; ++k) o[9 + k] = (double)(dl[k] / lam + di[k]);
        if (abs_out) {
            /* magnitude scale of each term: used to state errors of cancelling sums */
            double *a = abs_out + (i - row_begin) * 7;
            a[0] = (double)(FABS(LOG(lam)) + FABS(theta * (eT - 1)) + FABS(mu0 * (phiA - phiB)));
            for (int k = 0; k < 6; ++k) a[1 + k] = (double)(FABS(dl[k] / lam) + FABS(di[k]));
        }
    }
}

/*
 * Whole-data result: ll, gradient[6], and the matching magnitude scales
 * scale[0] = sum_i |terms of ll_i| + |constant|, scale[1+k] = sum_i |terms of g_ik|.
 * probs (may be NULL) receives p_i for all i.
 */
static int FN(oracle_total)(int D, long N, const double *locs, const double *times, const double *params,
                            double *ll_out, double *grad_out, double *scale_out, double *probs) {
    double *rows = (double *)malloc(sizeof(double) * 15 * (size_t)N);
    double *absr = (double *)malloc(sizeof(double) * 7 * (size_t)N);
    if (!rows || !absr) {
        free(rows);
        free(absr);
        return -1;
    }
    FN(oracle_rows)(D, N, locs, times, params, 0, N, rows, absr);
    FN(nsum_t) ll = {0, 0}, g[6] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    FN(nsum_t) sc[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (long i = 0; i < N; ++i) {
        FN(nsum_add)(&ll, rows[i * 15 + 7]);
        for (int k = 0; k < 6; ++k) FN(nsum_add)(&g[k], rows[i * 15 + 9 + k]);
        for (int k = 0; k < 7; ++k) FN(nsum_add)(&sc[k], absr[i * 7 + k]);
        if (probs) probs[i] = rows[i * 15 + 8];
    }
    const REAL cst = (REAL)N * (REAL)(D - 1) * LOG(INV_SQRT_2PI);
    FN(nsum_add)(&ll, cst);
    FN(nsum_add)(&sc[0], FABS(cst));
    if (ll_out) *ll_out = (double)FN(nsum_get)(&ll);
    if (grad_out)
        for (int k = 0; k < 6; ++k) grad_out[k] = (double)FN(nsum_get)(&g[k]);
    if (scale_out)
        for (int k = 0; k < 7; ++k) scale_out[k] = (double)FN(nsum_get)(&sc[k]);
    free(rows);
    free(absr);
    return 0;
}

/*
 * Location gradient out[k*D + d] = d ll / d x_kd. The reference has none (the JNI header
 * declares getLocationGradient, src/jni/dr_inference_hawkes_NativeHPHSingleton.h:92-96, and
 * the .cpp leaves it commented out, :80-89); this is the exact derivative of ll above. The
 * integral terms do not depend on locations, and with w_i = 1/lambda_i,
 * d b_ij/d x_i = -tauX^2 (x_i - x_j) b_ij, d s_ij/d x_i = -sigma^2 (x_i - x_j) s_ij:
 *   d ll/d x_k = - sum_j (x_k - x_j) [cB tauX^2 b_kj (w_k + w_j) + cS sigma^2 (s_kj w_k + s_jk w_j)].
 * scale_out (may be NULL) receives sum_j |terms| per component (the magnitude against which
 * errors of a cancelling sum are stated).
 */
static int FN(oracle_locgrad)(int D, long N, const double *locs, const double *times, const double *params,
                              double *out, double *scale_out) {
    const REAL sigma = params[0], tauX = params[1], tauT = params[2];
    const REAL omega = params[3], theta = params[4], mu0 = params[5];
    const REAL cB = mu0 * POW(tauX, (REAL)D) * tauT;
    const REAL cS = POW(sigma, (REAL)D) * theta * omega;
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    if (!w) return -1;
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < N; ++i) {
        FN(nsum_t) sb = {0, 0}, ss = {0, 0};
        const REAL ti = times[i];
        for (long j = 0; j < N; ++j) {
            REAL r2 = 0;
            for (int d = 0; d < D; ++d) {
                const REAL dx = (REAL)locs[i * D + d] - (REAL)locs[j * D + d];
                r2 += dx * dx;
            }
            const REAL r = SQRT(r2);
            const REAL dt = ti - (REAL)times[j];
            FN(nsum_add)(&sb, FN(pdf)(tauX * r) * FN(pdf)(tauT * dt));
            if (dt > 0) FN(nsum_add)(&ss, EXP(-omega * dt) * FN(pdf)(sigma * r));
        }
        w[i] = 1 / (cB * FN(nsum_get)(&sb) + cS * FN(nsum_get)(&ss));
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (long k = 0; k < N; ++k) {
        FN(nsum_t) g[8], a[8];
        for (int d = 0; d < D; ++d) g[d].s = g[d].c = a[d].s = a[d].c = 0;
        const REAL tk = times[k];
        for (long j = 0; j < N; ++j) {
            REAL dx[8], r2 = 0;
            for (int d = 0; d < D; ++d) {
                dx[d] = (REAL)locs[k * D + d] - (REAL)locs[j * D + d];
                r2 += dx[d] * dx[d];
            }
            const REAL r = SQRT(r2);
            const REAL dt = tk - (REAL)times[j];
            const REAL b = FN(pdf)(tauX * r) * FN(pdf)(tauT * dt);
            REAL f = cB * tauX * tauX * b * (w[k] + w[j]);
            if (dt > 0) f += cS * sigma * sigma * EXP(-omega * dt) * FN(pdf)(sigma * r) * w[k];
            if (dt < 0) f += cS * sigma * sigma * EXP(omega * dt) * FN(pdf)(sigma * r) * w[j];
            for (int d = 0; d < D; ++d) {
                FN(nsum_add)(&g[d], -f * dx[d]);
                FN(nsum_add)(&a[d], FABS(f * dx[d]));
            }
        }
        for (int d = 0; d < D; ++d) {
            out[k * D + d] = (double)FN(nsum_get)(&g[d]);
            if (scale_out) scale_out[k * D + d] = (double)FN(nsum_get)(&a[d]);
        }
    }
    free(w);
    return 0;
}

#undef FN
#undef CAT
#undef CAT_
