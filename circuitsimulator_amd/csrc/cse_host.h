/* cse_host.h -- host-side helpers of the engine (plain C, no HIP). */
#ifndef CSE_HOST_H
#define CSE_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11) and the stream convention shared with
 * the device: 32-bit word w of stream (key, id, domain) is output (w & 3) of
 * the block at counter {w >> 2, id lo, id hi, domain}. */
void cse_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t cse_rand64(uint64_t key, uint64_t id, uint32_t domain, uint64_t w);

/* thr[d] = floor(cdf(d) * 2^53), d = 0..n, cdf of eq. 79 (stabilizer.c:677-716) */
void cse_dimension_thresholds(int n, uint64_t *thr);

/* sampled estimators (innerprod.c:83,142) from per-sample exact Z[w] sums
 * z[4*s..] (scale 2^32) and exponents sexp[s] (-1: vanished), and from the
 * per-sample FP32-Kahan sums kz[2*s..] */
void cse_sampled_estimators(int t, int64_t count, const long long *z, const int32_t *sexp, const double *kz,
                            double *mean, double *mean_kahan);

/* The same in parts, for splitting the samples over threads: each part
 * holds the exact sums of A and B per projection exponent over a range of
 * samples and a compensated sum of the FP32-Kahan estimates.  The exact
 * sums of the parts add exactly; the compensated sums are added in part
 * order, so the Kahan mean depends on the split, never on the threads. */
typedef struct {
    __int128 ahi[65], bhi[65];
    unsigned __int128 alo[65], blo[65];
    double kacc, kcomp;
} cse_est_part;
void cse_est_accumulate(cse_est_part *part, int t, int64_t lo, int64_t hi, const long long *z,
                        const int32_t *sexp, const double *kz);
void cse_est_finish(const cse_est_part *parts, int nparts, int t, int64_t count, double *mean,
                    double *mean_kahan);

/* GF(2) rank of k rows of t bits (matrix.c:676-724 semantics) */
int cse_gf2_rank(const uint64_t *rows, int k);

/* double-double helpers */
typedef struct {
    double hi, lo;
} cse_dd;
cse_dd cse_dd_add(cse_dd a, cse_dd b);
cse_dd cse_dd_mul(cse_dd a, cse_dd b);
cse_dd cse_dd_from_i128(__int128 v);

#ifdef __cplusplus
}
#endif
#endif
