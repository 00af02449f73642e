/*
 * ref_harness.c -- test-only bridge to the reference's own C sources.
 *
 * Compiled (by oracle/Makefile, target `ref`) together with the reference's
 * libcirc/stabilizer/stabilizer.c, libcirc/utils/matrix.c and
 * libcirc/stateprep.c taken where they lie under the reference checkout; the
 * result goes to oracle/_ref/libref.so and is never committed.  It converts
 * between the oracle's uint64-row state (cse_oracle.h) and the reference's
 * struct StabilizerState (stabilizer.h:5-16) and calls the reference
 * functions unchanged.  Used only by tests/ and bench.py's cpu_baseline leg.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "stabilizer/stabilizer.h"
#include "cse_oracle.h"

/* defined in libcirc/stateprep.c (declared the same way in innerprod.c:9-10) */
struct StabilizerState *prepH(int i, int t);
struct StabilizerState *prepL(int i, int t, struct BitMatrix *L);

static struct StabilizerState *to_ref(const orc_state *s)
{
    struct StabilizerState *r = allocStabilizerState(s->n, s->k);
    r->Q = s->Q;
    for (int i = 0; i < s->n; i++) {
        BitVectorSet(r->h, i, (s->h >> i) & 1);
        BitVectorSet(r->D1, i, (s->D1 >> i) & 1);
        BitVectorSet(r->D2, i, (s->D2 >> i) & 1);
        for (int j = 0; j < s->n; j++) {
            BitMatrixSet(r->G, i, j, (s->G[i] >> j) & 1);
            BitMatrixSet(r->Gbar, i, j, (s->Gb[i] >> j) & 1);
            BitMatrixSet(r->J, i, j, (s->J[i] >> j) & 1);
        }
    }
    return r;
}

static void from_ref(const struct StabilizerState *r, orc_state *s)
{
    memset(s, 0, sizeof(*s));
    s->n = r->n;
    s->k = r->k;
    s->Q = r->Q;
    for (int i = 0; i < r->n; i++) {
        s->h |= (uint64_t)BitVectorGet(r->h, i) << i;
        s->D1 |= (uint64_t)BitVectorGet(r->D1, i) << i;
        s->D2 |= (uint64_t)BitVectorGet(r->D2, i) << i;
        for (int j = 0; j < r->n; j++) {
            s->G[i] |= (uint64_t)BitMatrixGet(r->G, i, j) << j;
            s->Gb[i] |= (uint64_t)BitMatrixGet(r->Gbar, i, j) << j;
            s->J[i] |= (uint64_t)BitMatrixGet(r->J, i, j) << j;
        }
    }
}

static struct BitVector *vec_ref(uint64_t v, int n)
{
    struct BitVector *b = newBitVector(n);
    for (int i = 0; i < n; i++) BitVectorSet(b, i, (v >> i) & 1);
    return b;
}

void ref_innerProduct(const orc_state *s1, const orc_state *s2, int *eps, int *p, int *m)
{
    struct StabilizerState *a = to_ref(s1), *b = to_ref(s2);
    innerProductExact(a, b, eps, p, m);
    freeStabilizerState(a);
    freeStabilizerState(b);
}

void ref_exponentialSum(orc_state *s, int *eps, int *p, int *m)
{
    struct StabilizerState *a = to_ref(s);
    exponentialSumExact(a, eps, p, m);
    from_ref(a, s);
    freeStabilizerState(a);
}

int ref_shrink(orc_state *s, uint64_t xi, int alpha, int lazy)
{
    struct StabilizerState *a = to_ref(s);
    struct BitVector *x = vec_ref(xi, s->n);
    int r = shrink(a, x, alpha, lazy);
    from_ref(a, s);
    BitVectorFree(x);
    freeStabilizerState(a);
    return r;
}

double ref_measurePauli(orc_state *s, int m, uint64_t zeta, uint64_t xi)
{
    struct StabilizerState *a = to_ref(s);
    struct BitVector *z = vec_ref(zeta, s->n), *x = vec_ref(xi, s->n);
    double r = measurePauli(a, m, z, x);
    from_ref(a, s);
    BitVectorFree(z);
    BitVectorFree(x);
    freeStabilizerState(a);
    return r;
}

void ref_prepH(orc_state *s, int i, int t)
{
    struct StabilizerState *a = prepH(i, t);
    from_ref(a, s);
    freeStabilizerState(a);
}

static struct BitMatrix *mat_ref(const uint64_t *L, int k, int t)
{
    struct BitMatrix *M = newBitMatrixZero(k, t);
    for (int i = 0; i < k; i++)
        for (int j = 0; j < t; j++) BitMatrixSet(M, i, j, (L[i] >> j) & 1);
    return M;
}

void ref_prepL(orc_state *s, int i, int t, const uint64_t *L, int k)
{
    struct BitMatrix *M = mat_ref(L, k, t);
    struct StabilizerState *a = prepL(i, t, M);
    from_ref(a, s);
    freeStabilizerState(a);
    BitMatrixFree(M);
}

/*
 * Timed CPU baseline of the sampled-mode inner loop (innerprod.c:119-139):
 * for each of the ntheta thetas and each i < nphi (<= 2^k),
 * innerProduct(theta', prepL(i)) with the reference's own functions, with
 * nthreads POSIX threads splitting the theta range cyclically like the
 * reference's rank-strided loop (probability.c:268).  The projected thetas
 * are supplied by the caller in the oracle layout.  Returns the number of
 * inner products computed; per-theta exact sums go to csum[4*theta], and
 * (ref_sampled_sweep_terms) every term to terms[theta * nphi + i].
 */
typedef struct {
    const orc_state *thetas;
    int ntheta, t, k, nthreads, tid, E, nphi;
    const uint64_t *L;
    int64_t *csum;
    int32_t *terms;
    long count;
} ref_job;

static void ref_zomega_add(int64_t c[4], int E, int eps, int p, int m);

static void *ref_worker(void *arg)
{
    ref_job *j = (ref_job *)arg;
    struct BitMatrix *M = mat_ref(j->L, j->k, j->t);
    for (int th = j->tid; th < j->ntheta; th += j->nthreads) {
        struct StabilizerState *a = to_ref(&j->thetas[th]);
        int64_t c[4] = {0, 0, 0, 0};
        for (int i = 0; i < j->nphi; i++) {
            struct StabilizerState *phi = prepL(i, j->t, M);
            int eps, p, m;
            innerProductExact(a, phi, &eps, &p, &m);
            ref_zomega_add(c, j->E, eps, p, m);
            if (j->terms)
                j->terms[(long)th * j->nphi + i] = eps ? (-p) * 8 + (((m % 8) + 8) % 8) + 1 : 0;
            freeStabilizerState(phi);
            j->count++;
        }
        memcpy(&j->csum[4 * th], c, sizeof(c));
        freeStabilizerState(a);
    }
    BitMatrixFree(M);
    return NULL;
}

/* Z[w] accumulation (the oracle's orc_zomega_add), kept local so libref.so stands alone */
static void ref_zomega_add(int64_t c[4], int E, int eps, int p, int m)
{
    if (!eps) return;
    int e = E + (p >> 1);
    int64_t v = (int64_t)1 << e;
    int mm = ((m % 8) + 8) % 8;
    int q0 = (p & 1) ? mm + 1 : mm;
    int qs[2] = {q0 % 8, (mm + 3) % 8};
    int sg[2] = {1, -1};
    int nt = (p & 1) ? 2 : 1;
    for (int u = 0; u < nt; u++) {
        int q = qs[u], s = sg[u];
        if (q >= 4) { q -= 4; s = -s; }
        c[q] += s * v;
    }
}

/* as ref_sampled_sweep, and every term packed as 0 (eps = 0) or
 * (-p) * 8 + (m mod 8) + 1 at terms[theta * nphi + i] when terms != NULL */
long ref_sampled_sweep_terms(const orc_state *thetas, int ntheta, int t, const uint64_t *L, int k,
                             int E, int nthreads, int nphi, int64_t *csum, int32_t *terms)
{
    if (nphi > (1 << k) || nphi < 0) nphi = 1 << k;
    if (nthreads < 1) nthreads = 1;
    pthread_t th[256];
    ref_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int i = 0; i < nthreads; i++) {
        jobs[i] = (ref_job){thetas, ntheta, t, k, nthreads, i, E, nphi, L, csum, terms, 0};
        pthread_create(&th[i], NULL, ref_worker, &jobs[i]);
    }
    long total = 0;
    for (int i = 0; i < nthreads; i++) {
        pthread_join(th[i], NULL);
        total += jobs[i].count;
    }
    return total;
}

long ref_sampled_sweep(const orc_state *thetas, int ntheta, int t, const uint64_t *L, int k,
                       int E, int nthreads, int nphi, int64_t *csum)
{
    return ref_sampled_sweep_terms(thetas, ntheta, t, L, k, E, nthreads, nphi, csum, NULL);
}
