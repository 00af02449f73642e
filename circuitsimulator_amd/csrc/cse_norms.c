/*
 * cse_norms.c -- the backend's per-projector computation, shared by the
 * drop-in executable and the Python front end: ||P|H^t>||^2 exactly
 * (innerprod.c:148-199) or as a median of means of sampled estimates
 * (innerprod.c:23-84), with the empty-projector and t = 0 closed forms, and
 * the sampling-disable rule of probability.c:162-174.  The engine context is
 * created on the first call that needs the GPU.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cse.h"

/* The context replays its batched calls from captured HIP graphs
 * (cse_set_graphs): repeated probability evaluations and sampling rounds
 * have the same shape call after call.  CSE_GRAPHS=0 in the environment
 * keeps them eager. */
static int ensure_ctx(cse_ctx **ctx, int device)
{
    if (*ctx) return CSE_OK;
    int rc = cse_create(device, ctx);
    if (rc != CSE_OK) return CSE_ERR_NODEV;
    const char *g = getenv("CSE_GRAPHS");
    return cse_set_graphs(*ctx, !(g && strcmp(g, "0") == 0));
}

/* 1 + #{+I} - #{-I} over the generators (innerprod.c:157-167) */
static double clifford_sum(const cse_projector *P)
{
    double sum = 1;
    for (int g = 0; g < P->nstabs; g++) {
        if (P->phase[g] == 0) sum += 1;
        if (P->phase[g] == 2) sum -= 1;
    }
    return sum;
}

/* samples per engine call (about 1 GiB of device records) */
#define CSE_CALL_SAMPLES (1 << 20)

/* mean over samples [first, first + count), in calls of at most
 * CSE_CALL_SAMPLES; a range that fits one call is that call's mean */
static int sampled_mean(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k, int exact,
                        int64_t count, uint64_t seed, int64_t first, double *out)
{
    double total = 0;
    const int64_t step = CSE_CALL_SAMPLES;
    for (int64_t s0 = 0; s0 < count; s0 += step) {
        const int64_t cnt = count - s0 < step ? count - s0 : step;
        cse_sampled_result r;
        memset(&r, 0, sizeof(r));
        const int rc = cse_sampled_projector(ctx, P, L, k, exact, seed, first + s0, cnt, 0, &r);
        if (rc != CSE_OK) return rc;
        if (cnt == count) {
            *out = r.mean;
            return CSE_OK;
        }
        total += r.mean * (double)cnt;
    }
    *out = total / (double)count;
    return CSE_OK;
}

/* vals[b] = mean over bin b's samples [b * samples, (b + 1) * samples): bins
 * that fit one call together go through cse_batch_sampled, a bin per
 * instance; larger bins are taken one at a time */
static int bin_means(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k, int exact, int samples,
                     int bins, uint64_t seed, double *vals)
{
    if (samples > CSE_CALL_SAMPLES / 2) {
        for (int b = 0; b < bins; b++) {
            const int rc = sampled_mean(ctx, P, L, k, exact, samples, seed, (int64_t)b * samples, &vals[b]);
            if (rc != CSE_OK) return rc;
        }
        return CSE_OK;
    }
    const int per = CSE_CALL_SAMPLES / samples;
    const int n = bins < per ? bins : per;
    cse_sampled_instance *in = (cse_sampled_instance *)calloc((size_t)n, sizeof(cse_sampled_instance));
    cse_sampled_result *res = (cse_sampled_result *)calloc((size_t)n, sizeof(cse_sampled_result));
    int rc = (in && res) ? CSE_OK : CSE_ERR_NOMEM;
    for (int i = 0; i < n && rc == CSE_OK; i++) {
        in[i].P = *P;
        in[i].exact = exact;
        in[i].k = exact ? 0 : k;
        for (int r = 0; !exact && r < k; r++) in[i].L[r] = L[r];
        in[i].seed = seed;
        in[i].count = samples;
    }
    for (int b0 = 0; b0 < bins && rc == CSE_OK; b0 += n) {
        const int m = bins - b0 < n ? bins - b0 : n;
        for (int i = 0; i < m; i++) in[i].first = (int64_t)(b0 + i) * samples;
        memset(res, 0, sizeof(cse_sampled_result) * (size_t)m);
        rc = cse_batch_sampled(ctx, in, m, 0, res, NULL);
        for (int i = 0; rc == CSE_OK && i < m; i++) vals[b0 + i] = res[i].mean;
    }
    free(in);
    free(res);
    return rc;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int cse_sampling_kept(int noapprox, int forceSample, int exact, int t, int k, int samples, int bins)
{
    if (noapprox) return 0;
    if (forceSample) return 1;
    const double lim = exact ? pow(2, (t + 1) / 2) - 1 : pow(2, k) - 1;
    return (double)samples * (double)bins * 2 > lim ? 0 : 1;
}

/* the cases that need no engine: 1 with *value set, else 0 */
static int closed_form(const cse_projector *P, double norm, int sampled, int samples, int bins, double *value)
{
    if (P->nstabs == 0) {
        *value = norm * norm;
        return 1;
    }
    if (P->nqubits == 0) {
        /* the sampled form keeps the norm factor, the exact one does not */
        const double f = clifford_sum(P) / (1 + (double)P->nstabs);
        *value = sampled ? norm * norm * f : f;
        return 1;
    }
    if (sampled && (samples <= 0 || bins <= 0)) {
        /* the reference divides 0 by 0 here (innerprod.c:83) */
        *value = NAN;
        return 1;
    }
    return 0;
}

int cse_projector_norm(cse_ctx **ctx, int device, const cse_projector *P, const uint64_t *L, int k, int exact,
                       double norm, int sampled, int samples, int bins, uint64_t seed, double *value)
{
    if (!ctx || !P || !value) return CSE_ERR_ARG;
    if (closed_form(P, norm, sampled, samples, bins, value)) return CSE_OK;
    int rc = ensure_ctx(ctx, device);
    if (rc != CSE_OK) return rc;
    if (!sampled) {
        cse_exact_result r;
        memset(&r, 0, sizeof(r));
        rc = cse_exact_projector(*ctx, P, L, k, exact, &r);
        if (rc == CSE_OK) *value = r.value;
        return rc;
    }
    double *vals = (double *)malloc(sizeof(double) * (size_t)bins);
    if (!vals) return CSE_ERR_NOMEM;
    rc = bin_means(*ctx, P, L, k, exact, samples, bins, seed, vals);
    if (rc == CSE_OK) {
        qsort(vals, (size_t)bins, sizeof(double), cmp_double);
        *value = (bins % 2) ? vals[(bins - 1) / 2] : (vals[bins / 2] + vals[bins / 2 - 1]) / 2;
    }
    free(vals);
    return rc;
}

int cse_projector_norms(cse_ctx **ctx, int device, int nproj, const cse_projector *P, const uint64_t *seeds,
                        const uint64_t *L, int k, int exact, double norm, int sampled, int samples, int bins,
                        double *values)
{
    if (!ctx || !P || !values || nproj < 0 || (sampled && nproj > 0 && !seeds)) return CSE_ERR_ARG;
    if (sampled) {
        for (int i = 0; i < nproj; i++) {
            const int rc = cse_projector_norm(ctx, device, &P[i], L, k, exact, norm, sampled, samples, bins,
                                              seeds[i], &values[i]);
            if (rc != CSE_OK) return rc;
        }
        return CSE_OK;
    }
    /* exact pair sums: every projector the closed forms leave is one
       instance of one batched launch */
    int *which = (int *)malloc(sizeof(int) * (size_t)(nproj > 0 ? nproj : 1));
    if (!which) return CSE_ERR_NOMEM;
    int m = 0;
    for (int i = 0; i < nproj; i++)
        if (!closed_form(&P[i], norm, 0, samples, bins, &values[i])) which[m++] = i;
    int rc = CSE_OK;
    if (m > 0) rc = ensure_ctx(ctx, device);
    cse_instance *in = NULL;
    cse_exact_result *res = NULL;
    if (rc == CSE_OK && m > 0) {
        in = (cse_instance *)calloc((size_t)m, sizeof(cse_instance));
        res = (cse_exact_result *)calloc((size_t)m, sizeof(cse_exact_result));
        if (!in || !res) rc = CSE_ERR_NOMEM;
    }
    if (rc == CSE_OK && m > 0) {
        for (int j = 0; j < m; j++) {
            in[j].P = P[which[j]];
            in[j].exact = exact ? 1 : 0;
            in[j].k = exact ? 0 : k;
            if (!exact)
                for (int r = 0; r < k; r++) in[j].L[r] = L[r];
        }
        rc = cse_batch_exact(*ctx, in, m, 0, res, NULL, NULL);
        if (rc == CSE_OK)
            for (int j = 0; j < m; j++) values[which[j]] = res[j].value;
    }
    free(in);
    free(res);
    free(which);
    return rc;
}
