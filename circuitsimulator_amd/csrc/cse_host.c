/*
 * cse_host.c -- host-side helpers of the engine (plain C): counter-based RNG,
 * the dimension law of random stabilizer states, exact-to-FP64 conversion of
 * the exact accumulators, and decompose (libcirc/probability.c:307-415).
 */
#include <math.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include "../../include/cse.h"
#include "cse_host.h"

/* ------------------------------------------------------------- Philox */
void cse_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static uint32_t rand32(uint64_t key, uint64_t id, uint32_t domain, uint64_t w)
{
    uint32_t ctr[4] = {(uint32_t)(w >> 2), (uint32_t)id, (uint32_t)(id >> 32), domain};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    uint32_t out[4];
    cse_philox4x32_10(ctr, k, out);
    return out[w & 3];
}

uint64_t cse_rand64(uint64_t key, uint64_t id, uint32_t domain, uint64_t w)
{
    return (uint64_t)rand32(key, id, domain, w) | ((uint64_t)rand32(key, id, domain, w + 1) << 32);
}

/* ------------------------------------------------ dimension law, eq. 79 */
static double logeta(int d, int n)
{
    if (d == 0) return 0.0;
    double prod = 0;
    for (int a = 1; a <= d; a++) {
        prod += log2(1 - pow(2, d - n - a));
        prod -= log2(1 - pow(2, -a));
    }
    return (-d * (d + 1) / 2) + prod;
}

void cse_dimension_thresholds(int n, uint64_t *thr)
{
    double dist[CSE_NMAX + 1], cum[CSE_NMAX + 1];
    double sum = 0;
    for (int d = 0; d <= n; d++) {
        dist[d] = pow(2, logeta(d, n));
        sum += dist[d];
    }
    for (int d = 0; d <= n; d++) dist[d] /= sum;
    for (int i = 0; i <= n; i++) {
        cum[i] = 0;
        for (int d = 0; d <= i; d++) cum[i] += dist[d];
    }
    for (int i = 0; i <= n; i++) cum[i] /= cum[n];
    for (int i = 0; i <= n; i++) thr[i] = (uint64_t)floor(cum[i] * 9007199254740992.0);
}

/* ------------------------------------------------------- double-double */
static cse_dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    cse_dd r = {s, e};
    return r;
}

static cse_dd quick_two_sum(double a, double b)
{
    double s = a + b;
    cse_dd r = {s, b - (s - a)};
    return r;
}

cse_dd cse_dd_add(cse_dd a, cse_dd b)
{
    cse_dd s = two_sum(a.hi, b.hi);
    cse_dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

cse_dd cse_dd_mul(cse_dd a, cse_dd b)
{
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p, e);
}

cse_dd cse_dd_from_i128(__int128 v)
{
    double hi = (double)v;
    __int128 rest = v - (__int128)hi;
    cse_dd r = {hi, (double)rest};
    return quick_two_sum(r.hi, r.lo);
}

static cse_dd dd_scale(cse_dd a, int e)
{
    cse_dd r = {ldexp(a.hi, e), ldexp(a.lo, e)};
    return r;
}

/* sqrt(2) to double-double precision */
static const cse_dd SQRT2 = {1.4142135623730951, -9.667293313452913e-17};

/* ------------------------------------------------ exact -> FP64 */
void cse_hist_value(const int64_t *hist, int pmin, double *re, double *im)
{
    /* bin (P, M) = hist[8 (-P) + M], P in [pmin, 0]: count * 2^{P/2} w^M.
     * With 2^{P/2} = 2^{floor(P/2)} sqrt2^{P&1} and w^M = (cre + i cim) /
     * sqrt2^{M&1}, a row P puts its even-M bins at 2^{floor(P/2)} times
     * sqrt2^{P&1} and its odd-M bins at 2^{floor(P/2) - (1 - P&1)} times
     * sqrt2^{1 - P&1}.  The counts are summed exactly per (power of two,
     * sqrt2 or not, real or imaginary) in __int128 (at most 2^63 per bin),
     * and each non-empty sum is converted once. */
    static const int cre[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    static const int cim[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    const int rows = 1 - pmin;
    const int emin = (pmin >> 1) - 1; /* powers of two in [emin, 0] */
    enum { NE = 66 };
    if (emin < -(NE - 1)) {
        *re = *im = NAN;
        return;
    }
    __int128 acc[4][NE]; /* 0: re, 1: im, 2: re * sqrt2, 3: im * sqrt2 */
    unsigned char used[NE] = {0};
    for (int row = 0; row < rows; row++) {
        const int64_t *h = hist + 8 * row;
        if (!(h[0] | h[1] | h[2] | h[3] | h[4] | h[5] | h[6] | h[7])) continue;
        const int P = -row, podd = P & 1;
        /* even M: (c0 - c4) + (c2 - c6) i; odd M: (c1 - c3 - c5 + c7) + (c1 + c3 - c5 - c7) i */
        const __int128 er = (__int128)h[0] - h[4], ei = (__int128)h[2] - h[6];
        __int128 orr = 0, oi = 0;
        for (int M = 1; M < 8; M += 2) {
            orr += (__int128)h[M] * cre[M];
            oi += (__int128)h[M] * cim[M];
        }
        const int e0 = (P >> 1) - emin, q0 = podd ? 2 : 0;
        const int e1 = (P >> 1) - (podd ? 0 : 1) - emin, q1 = podd ? 0 : 2;
        if (!used[e0]) { used[e0] = 1; acc[0][e0] = acc[1][e0] = acc[2][e0] = acc[3][e0] = 0; }
        if (!used[e1]) { used[e1] = 1; acc[0][e1] = acc[1][e1] = acc[2][e1] = acc[3][e1] = 0; }
        acc[q0][e0] += er;
        acc[q0 + 1][e0] += ei;
        acc[q1][e1] += orr;
        acc[q1 + 1][e1] += oi;
    }
    cse_dd s[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < NE; i++) {
        if (!used[i]) continue;
        for (int q = 0; q < 4; q++)
            if (acc[q][i]) s[q] = cse_dd_add(s[q], dd_scale(cse_dd_from_i128(acc[q][i]), i + emin));
    }
    cse_dd r = cse_dd_add(s[0], cse_dd_mul(s[2], SQRT2));
    cse_dd im_ = cse_dd_add(s[1], cse_dd_mul(s[3], SQRT2));
    *re = r.hi + r.lo;
    *im = im_.hi + im_.lo;
}

/* (hi * 2^64 + lo) * 2^e in double-double; the carry of lo joins hi
 * exactly, leaving two exact 128-bit integers to convert */
static cse_dd dd_split_sum(__int128 hi, unsigned __int128 lo, int e)
{
    const __int128 top = hi + (__int128)(lo >> 64);
    const __int128 low = (__int128)(uint64_t)lo;
    return cse_dd_add(dd_scale(cse_dd_from_i128(top), e + 64), dd_scale(cse_dd_from_i128(low), e));
}

void cse_est_accumulate(cse_est_part *part, int t, int64_t lo, int64_t hi, const long long *z,
                        const int32_t *sexp, const double *kz)
{
    /* exact sums of A and B per projection exponent s (0..64): each
     * __int128 term is split at 2^64, the high parts summed signed and the
     * low parts unsigned (no overflow below 2^63 samples) */
    memset(part, 0, sizeof(*part));
    /* 2^(t - s) for t, s in [0, 64]: a product with a power of two is
     * exact, so this equals ldexp */
    double p2[129];
    for (int e = 0; e <= 128; e++) p2[e] = ldexp(1.0, e - 64);
    double kacc = 0, kcomp = 0;
    for (int64_t s = lo; s < hi; s++) {
        const int se = sexp[s];
        if (se < 0 || se > 64) continue;
        const __int128 c0 = z[4 * s], c1 = z[4 * s + 1], c2 = z[4 * s + 2], c3 = z[4 * s + 3];
        const __int128 A = c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3;
        const __int128 B = c0 * c1 - c0 * c3 + c1 * c2 + c2 * c3;
        part->ahi[se] += A >> 64;
        part->alo[se] += (uint64_t)A;
        part->bhi[se] += B >> 64;
        part->blo[se] += (uint64_t)B;
        /* FP32-Kahan path: |S|^2 of the FP32 sums, summed compensated in FP64 */
        double kr = kz[2 * s], ki = kz[2 * s + 1];
        double v = (kr * kr + ki * ki) * p2[t - se + 64];
        double yk = v - kcomp;
        double tk = kacc + yk;
        kcomp = (tk - kacc) - yk;
        kacc = tk;
    }
    part->kacc = kacc;
    part->kcomp = kcomp;
}

void cse_est_finish(const cse_est_part *parts, int nparts, int t, int64_t count, double *mean,
                    double *mean_kahan)
{
    cse_dd X = {0, 0}, Y = {0, 0};
    for (int se = 0; se <= 64; se++) {
        __int128 ahi = 0, bhi = 0;
        unsigned __int128 alo = 0, blo = 0;
        for (int i = 0; i < nparts; i++) {
            ahi += parts[i].ahi[se];
            alo += parts[i].alo[se];
            bhi += parts[i].bhi[se];
            blo += parts[i].blo[se];
        }
        const int e = t - se - 64;
        X = cse_dd_add(X, dd_split_sum(ahi, alo, e));
        Y = cse_dd_add(Y, dd_split_sum(bhi, blo, e));
    }
    cse_dd tot = cse_dd_add(X, cse_dd_mul(Y, SQRT2));
    /* division by count: count is an integer < 2^53, refine once */
    double q = (tot.hi + tot.lo) / (double)count;
    cse_dd prod = cse_dd_mul((cse_dd){q, 0}, (cse_dd){(double)count, 0});
    cse_dd rem = cse_dd_add(tot, (cse_dd){-prod.hi, -prod.lo});
    q += (rem.hi + rem.lo) / (double)count;
    *mean = q;
    /* the parts' compensated sums, added compensated in part order */
    double kacc = parts[0].kacc, kcomp = parts[0].kcomp;
    for (int i = 1; i < nparts; i++) {
        double yk = (parts[i].kacc - parts[i].kcomp) - kcomp;
        double tk = kacc + yk;
        kcomp = (tk - kacc) - yk;
        kacc = tk;
    }
    *mean_kahan = kacc / (double)count;
}

void cse_sampled_estimators(int t, int64_t count, const long long *z, const int32_t *sexp, const double *kz,
                            double *mean, double *mean_kahan)
{
    cse_est_part part;
    cse_est_accumulate(&part, t, 0, count, z, sexp, kz);
    cse_est_finish(&part, 1, t, count, mean, mean_kahan);
}

/* ------------------------------------------------------------- rank */
int cse_gf2_rank(const uint64_t *rows, int k)
{
    uint64_t r[CSE_NMAX];
    int n = 0;
    for (int i = 0; i < k && i < CSE_NMAX; i++) r[n++] = rows[i];
    int rank = 0;
    for (int col = 0; col < 64 && rank < n; col++) {
        int piv = -1;
        for (int i = rank; i < n; i++)
            if ((r[i] >> col) & 1ull) { piv = i; break; }
        if (piv < 0) continue;
        uint64_t tmp = r[piv]; r[piv] = r[rank]; r[rank] = tmp;
        for (int i = 0; i < n; i++)
            if (i != rank && ((r[i] >> col) & 1ull)) r[i] ^= r[rank];
        rank++;
    }
    return rank;
}

/* ----------------------------------------------------------- decompose */
#define CSE_DOM_L 2u

/* #{i : |x_i| = w} over x_i = sum_j bit_{k-1-j}(i) L_j, walked in Gray-code
 * order (the multiset does not depend on the order) */
static void lspace_weights_host(const uint64_t *L, int k, int64_t *hist)
{
    memset(hist, 0, 65 * sizeof(int64_t));
    uint64_t x = 0;
    hist[0] = 1;
    for (uint64_t s = 1; s < (1ull << k); s++) {
        x ^= L[k - 1 - __builtin_ctzll(s)];
        hist[__builtin_popcountll(x)]++;
    }
}

/* Z_L = sum_x 2^-floor(|x|/2) from the weight histogram (the integer
 * division of probability.c:396), accumulated in long double and rounded
 * once */
static double fidelity_sum(const uint64_t *L, int k)
{
    int64_t hist[65];
    int done = 0;
    if (k >= 20) {
        int dev = 0;
        const char *s = getenv("CSE_DEVICE");
        if (s && *s) dev = atoi(s);
        cse_ctx *ctx = NULL;
        if (cse_create(dev, &ctx) == CSE_OK) {
            done = cse_lspace_weights(ctx, L, k, hist) == CSE_OK;
            cse_destroy(ctx);
        }
    }
    if (!done) lspace_weights_host(L, k, hist);
    long double z = 0;
    for (int w = 64; w >= 0; w--) z += ldexpl((long double)hist[w], -(w / 2));
    return (double)z;
}

int cse_decompose(int t, uint64_t seed, uint64_t *L, double *norm, int *exact, int *k, double fidbound,
                  int rank, int fidelity, int forceL, int verbose, int quiet)
{
    if (!L || !norm || !exact || !k) return CSE_ERR_ARG;
    if (t < 0 || t > CSE_NMAX) return CSE_ERR_ARG;
    if (t == 0) {
        *exact = 0;
        *norm = 1;
        return CSE_OK;
    }
    const double v = cos(M_PI / 8);
    *norm = pow(2, floor((float)t / 2) / 2);
    if (t % 2) *norm *= 2 * v;
    if (*exact) return CSE_OK;

    int forceK = 1;
    if (*k <= 0) {
        forceK = 0;
        *k = (int)ceil(1 - 2 * t * log2(v) - log2(fidbound));
        if (verbose) printf("Autopicking k = %d to achieve delta = %f.\n", *k, fidbound);
    }
    if (*k > t / 2 && !forceK && !forceL) {
        if (verbose) printf("k > t/2. Reverting to exact decomposition.\n");
        *exact = 1;
        return CSE_OK;
    }
    if (*k > t) {
        if (forceK && !quiet) printf("Can't have k > t. Setting k to %d.\n", t);
        *k = t;
    }
    if (*k > 40) return CSE_ERR_ARG;
    const uint64_t tm = t == 64 ? ~0ull : ((1ull << t) - 1);
    double innerProduct = 0, Z_L = 0;
    uint64_t draw = 0;
    uint64_t w = 0;
    while (innerProduct < 1 - fidbound || forceK) {
        for (int r = 0; r < *k; r++) {
            L[r] = cse_rand64(seed, draw, CSE_DOM_L, w) & tm;
            w += 2;
        }
        draw++;
        if (rank) {
            if (cse_gf2_rank(L, *k) < *k) {
                if (!quiet) printf("L has insufficient rank. Sampling again...\n");
                continue;
            }
        }
        if (fidelity) {
            Z_L = fidelity_sum(L, *k);
            innerProduct = pow(2, *k) * pow(v, 2 * t) / Z_L;
            if (forceK) {
                printf("delta = 1 - <H^t|L>: %lf\n", 1 - innerProduct);
                break;
            } else if (innerProduct < 1 - fidbound) {
                if (!quiet) printf("delta = 1 - <H^t|L>: %lf - Not good enough!\n", 1 - innerProduct);
            } else {
                if (!quiet) printf("delta = 1 - <H^t|L>: %lf\n", 1 - innerProduct);
            }
        } else {
            break;
        }
    }
    if (fidelity) *norm = sqrt(pow(2, *k) * Z_L);
    return CSE_OK;
}

int cse_decompose_many(int t, const uint64_t *seeds, int n, uint64_t *L, int ldl, double *norm, int *exact, int *k,
                       double fidbound, int rank, int fidelity, int forceL, int verbose, int quiet)
{
    if (!seeds || !L || !norm || !exact || !k || n < 0 || ldl < 1) return CSE_ERR_ARG;
    const int exact0 = *exact, k0 = *k;
    for (int i = 0; i < n; i++) {
        uint64_t rows[CSE_NMAX];
        *exact = exact0;
        *k = k0;
        const int rc = cse_decompose(t, seeds[i], rows, &norm[i], exact, k, fidbound, rank, fidelity, forceL,
                                     verbose, quiet);
        if (rc != CSE_OK) return rc;
        const int nr = *exact ? 0 : *k;
        if (nr > ldl) return CSE_ERR_ARG;
        for (int r = 0; r < nr; r++) L[(size_t)i * (size_t)ldl + (size_t)r] = rows[r];
    }
    return CSE_OK;
}
