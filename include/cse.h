/*
 * cse.h -- C ABI of the MI355X (gfx950) stabilizer-rank inner-product engine
 * (libcse.so).  Plain pointers and sizes; every buffer is a host buffer owned
 * by the caller; the engine owns its device memory.  One context per host
 * thread; every call is synchronous on the context's HIP stream.
 *
 * Return values: CSE_OK (0) or a negative CSE_ERR_* code; cse_error() gives
 * the message of the last failure on the context, cse_error(NULL) that of the
 * calling thread's last failed cse_create.
 */
#ifndef CSE_H
#define CSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSE_OK 0
#define CSE_ERR_ARG (-1)
#define CSE_ERR_HIP (-2)
#define CSE_ERR_NOMEM (-3)
#define CSE_ERR_RNG (-4)
#define CSE_ERR_NODEV (-5)   /* no usable gfx950 device for a context */

#define CSE_NMAX 64

/*
 * Packed stabilizer state (semantics of libcirc/stabilizer/stabilizer.h:5-16),
 * CSE_STATE_WORDS(n) little-endian uint64 words:
 *   w[0] = n | k << 8 | Q << 16        (n <= 64, 0 <= k <= n, Q in Z8)
 *   w[1] = h        bit j = coordinate j
 *   w[2] = D1, w[3] = D2               D_a = 2*D1_a + 4*D2_a
 *   w[4 .. 4+n)      rows of G         bit j of row a = G[a][j]
 *   w[4+n .. 4+2n)   rows of Gbar = (G^-1)^T
 *   w[4+2n .. 4+3n)  rows of J         J_ab in {0,4} stored as a bit
 */
#define CSE_STATE_WORDS(n) (3 * (n) + 4)

typedef struct cse_ctx cse_ctx;

/* Projector = product of (1 + g)/2 over r commuting Hermitian Paulis
 * g = i^phase Z(zs) X(xs); mirrors struct Projector (libcirc/utils/comms.h:4-11)
 * with phase = 2*phaseSign + phaseComplex (innerprod.c:101-102). */
typedef struct {
    int nstabs;
    int nqubits;
    int phase[CSE_NMAX];
    uint64_t xs[CSE_NMAX];
    uint64_t zs[CSE_NMAX];
} cse_projector;

/* A context on a gfx950 device.  It reads CSE_FLUSH_TERMS from the
 * environment when set: the number of terms of one instance after which a
 * wave of the exact-mode sweep flushes its partial histogram (1 to 2^30;
 * default 2^30, which keeps the sweep's 32-bit partial bins from wrapping).
 * Results do not depend on it; a value outside that range fails the call. */
int cse_create(int device, cse_ctx **ctx);
void cse_destroy(cse_ctx *ctx);
const char *cse_error(const cse_ctx *ctx);
/* number of compute units and the device name, for reports */
int cse_device_info(cse_ctx *ctx, int *cus, char *name, int name_len);

/*
 * enable != 0: the context's batched calls (cse_exact_projector,
 * cse_sampled_projector, cse_batch_exact, cse_batch_sampled and the
 * projector norms) replay HIP graphs: the first call of a given shape
 * (instance count, records, work items, buffers) captures its plan copy,
 * clearing launch, record launch, sweep and results copy, and later calls
 * of that shape launch the captured graph after writing their plan into
 * the same pinned stage.  Results are identical; calls that return
 * per-term output always run eagerly.  No reference counterpart: the
 * reference starts one backend process per probability() call
 * (probability.py:237-244).  Default off.
 */
int cse_set_graphs(cse_ctx *ctx, int enable);

/*
 * The sweep kernel of the context's calls of one mode, a tuning knob (the
 * results are identical).  mode 0 (sampled calls): kernel 0 the engine's
 * choice (2 when every instance of the call uses the |H^t> decomposition,
 * else 1), 1 the eight-wave sweep whose hand-scheduled path takes records
 * of at most two non-pivot columns (others run compiled code), 2 the
 * eight-wave sweep whose hand-scheduled path takes any number of non-pivot
 * columns and, for |H^t> terms, retires uncoupled variables in one batch
 * and runs each aligned batch of 64 terms as eight groups of 8.  mode 1
 * (exact calls): kernel 0 the engine's choice (2 for launches of at least 16
 * slices of 64 j per resident wave, else 1), 1 the exact sweep at seven
 * waves per SIMD, 2 the same at eight.  No reference counterpart (the
 * reference has one CPU code path).
 */
int cse_set_sweep_kernel(cse_ctx *ctx, int mode, int kernel);

/*
 * <s2[i] | s1[i]> for npairs state pairs of dimension n: one wavefront per
 * pair.  Replaces innerProductExact (libcirc/stabilizer/stabilizer.h:29,
 * stabilizer.c:589-659).  out[3*i .. 3*i+2] = (eps, p, m) with the value
 * eps * 2^{p/2} * e^{i pi m/4}; p and m are meaningful only when eps = 1
 * (then m is reduced mod 8).  one_launch_per_pair != 0 issues one kernel
 * launch per pair (the plain kernel); otherwise one launch covers all pairs.
 * If ms is not NULL it receives the device time of the launches.
 */
int cse_inner_products(cse_ctx *ctx, int n, size_t npairs, const uint64_t *s1,
                       const uint64_t *s2, int32_t *out, int one_launch_per_pair, double *ms);

/*
 * Decomposition states, packed: out[c] = prepL(idx[c], t, L) when exact == 0
 * (libcirc/stateprep.c:85-120; L has k rows of t bits) or prepH(idx[c], t)
 * when exact != 0 (stateprep.c:36-81).  Built on the device.
 */
int cse_prep_states(cse_ctx *ctx, int exact, int t, const uint64_t *L, int k,
                    const uint64_t *idx, size_t count, uint64_t *out);

/*
 * Random stabilizer states theta_s for s in [first, first + count), drawn on
 * the device from Philox4x32-10 keyed by seed (the draw of
 * randomStabilizerState, stabilizer.c:689-756), then projected by P
 * (innerprod.c:95-116; P may be NULL).  out: packed states (may be NULL);
 * sexp[c] = number of 2^-1/2 factors of the projection, -1 if it vanished.
 */
int cse_random_states(cse_ctx *ctx, int n, uint64_t seed, int64_t first, size_t count,
                      const cse_projector *P, uint64_t *out, int32_t *sexp);

/* ---------------------------------------------------------------- results */
/* Exact-mode histograms: P = p - s runs over [CSE_HIST_PMIN, 0] (p >= -64 per
 * inner product, s <= 64 projection factors), 8 phases each. */
#define CSE_HIST_PMIN (-128)
#define CSE_HIST_BINS ((1 - CSE_HIST_PMIN) * 8)   /* 1032 */

typedef struct {
    /* exact accumulation: counts of terms 2^{P/2} e^{i pi M/4} per bin
     * hist[(-P)*8 + M], P in [pmin, 0]; off-diagonal terms 2 Re(z) are
     * counted as z + conj(z) (innerprod.c:254-260) */
    int pmin;           /* output: always CSE_HIST_PMIN */
    int64_t *hist;      /* caller buffer of CSE_HIST_BINS entries, or NULL */
    double re, im;      /* the exact total rounded once to FP64 */
    double value;       /* |total| -- exactProjector's return (innerprod.c:198) */
    double kahan_re, kahan_im; /* Kahan-compensated FP32 sums, reduced in FP64 */
    int64_t terms;      /* pair terms l evaluated, chi (chi + 1) / 2 */
    int64_t ips;        /* inner products computed (non-vanishing projections) */
    double kernel_ms;   /* device time of the pair sweep */
    double total_ms;    /* device time of the whole call */
} cse_exact_result;

/*
 * Exact projector norm sum_{i<=j} <phi_j | Pi | phi_i> with 2 Re on i < j:
 * replaces exactProjector / exactProjectorWork (innerprod.c:148-261) for
 * t >= 1 and a non-empty P (the empty and t = 0 cases are closed forms the
 * caller handles).  exact != 0: |H^t> decomposition (chi = 2^ceil(t/2));
 * else |L> from L (k rows, chi = 2^k).  chi <= 2^30 (CSE_ERR_ARG otherwise),
 * the range in which every sum stays exact in int64.
 */
int cse_exact_projector(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k,
                        int exact, cse_exact_result *res);

/*
 * cse_exact_projector without any allocation.  It has no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls
 * no hipMalloc, hipFree, host allocator or other allocating routine: every
 * device buffer the call needs (work queue, plan, results block, records)
 * lives in a static device array of 8,762,624 bytes that each GPU holds
 * until the process ends, and the host stages are on the stack.  Results
 * equal cse_exact_projector's.  chi <= 2^13 (|H^t> with t <= 26, or k <= 13);
 * larger chi returns CSE_ERR_ARG.  Every call on one GPU shares that GPU's
 * array, from any context: calls that overlap on one GPU race.
 */
int cse_exact_projector_noalloc(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k,
                                int exact, cse_exact_result *res);

typedef struct {
    double mean;        /* mean over samples of 2^t |pf * sum_i <phi_i|theta'>|^2,
                           from exact per-sample int64 sums (innerprod.c:45-144) */
    double mean_kahan;  /* the same estimator from the Kahan-FP32 per-sample sums */
    int64_t ips;        /* inner products computed */
    int64_t vanished;   /* samples whose projection vanished (term 0) */
    double kernel_ms;   /* device time of the inner-product sweep */
    double prep_ms;     /* device time of theta generation + projection */
    /* optional per-sample outputs (caller buffers or NULL) */
    int64_t *zsum;      /* 4 per sample: sum_i <phi_i|theta'> = 2^-32 (c0 + c1 w + c2 w^2 + c3 w^3) */
    int32_t *sexp;      /* projection exponent, -1 if vanished */
    double *kahan;      /* 2 per sample: FP32 Kahan re, im */
    int32_t *terms;     /* chi per sample: 0 if eps = 0, else (-p)*8 + m + 1 */
} cse_sampled_result;

/*
 * Sampled projector norm: mean over samples s in [first, first + count) of
 * 2^t |pf_s sum_i <phi_i | theta_s'>|^2 for Philox-drawn random stabilizer
 * states theta_s (replaces sampledProjector / singleProjectorSample,
 * innerprod.c:45-144; the median over bins of multiSampledProjector,
 * innerprod.c:23-41, is taken by the caller over disjoint sample ranges).
 * Persistent launch over a device work queue of (sample, chunk) items;
 * chunk = 0 picks the tuned default.
 */
int cse_sampled_projector(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k,
                          int exact, uint64_t seed, int64_t first, int64_t count, int chunk,
                          cse_sampled_result *res);

/* --------------------------------------------- batched uneven instances */
typedef struct {
    cse_projector P;
    int exact;          /* as in cse_exact_projector */
    int k;
    uint64_t L[CSE_NMAX];
} cse_instance;

/*
 * Many independent exact-mode instances with uneven (t, r, k), every pair term
 * of every instance in one persistent launch from a device work queue.
 * res[i] as in cse_exact_projector (hist may be NULL).  terms_out, if not
 * NULL, receives every term's packed (eps, p, m) as in cse_sampled_result.terms,
 * instance after instance, l-major (innerprod.c:211-215 order).
 */
int cse_batch_exact(cse_ctx *ctx, const cse_instance *inst, int ninst, int chunk,
                    cse_exact_result *res, int32_t *terms_out, double *kernel_ms);

typedef struct {
    cse_projector P;
    int exact;          /* as in cse_sampled_projector */
    int k;
    uint64_t L[CSE_NMAX];
    uint64_t seed;      /* Philox key of the random states */
    int64_t first;      /* sample range [first, first + count) */
    int64_t count;
} cse_sampled_instance;

/*
 * Many sampled-mode instances (the bins of multiSampledProjector,
 * innerprod.c:23-41, or the norms of many probability calls) in one record
 * launch and one persistent sweep.  res[i] is exactly what
 * cse_sampled_projector returns for instance i (its optional per-sample
 * buffers filled when non-NULL; the timings are those of the whole call).
 * At most 2^24 samples over all instances.
 */
int cse_batch_sampled(cse_ctx *ctx, const cse_sampled_instance *inst, int ninst, int chunk,
                      cse_sampled_result *res, double *kernel_ms);

/* --------------------------------------------------- host-side helpers */
/*
 * decompose (libcirc/probability.c:307-415): norm of the decomposition, the
 * exact/L choice and L itself (rows drawn from the seeded host RNG; rank and
 * fidelity checks as in the reference).  Returns CSE_OK; *exact and *k are
 * in/out as in the reference.  The fidelity's Z_L sums 2^-floor(|x|/2)
 * over the 2^k combinations from a histogram of |x| (cse_lspace_weights on
 * the GPU named by CSE_DEVICE when k >= 20 and one is present, a Gray-code
 * walk on the host otherwise), rounded once; the reference adds the 2^k
 * terms one by one in FP64.
 */
int cse_decompose(int t, uint64_t seed, uint64_t *L, double *norm, int *exact, int *k,
                  double fidbound, int rank, int fidelity, int forceL, int verbose, int quiet);

/*
 * cse_decompose once per seed, with the same t and options: the per-call
 * decompositions of a batch of probability evaluations (probability.c:
 * 307-415, one seed per call).  Seed i's rows go to L + i * ldl, its norm to
 * norm[i]; *exact and *k are in/out as in cse_decompose (they depend on t and
 * the options only).  CSE_ERR_ARG if k exceeds ldl.
 */
int cse_decompose_many(int t, const uint64_t *seeds, int n, uint64_t *L, int ldl, double *norm, int *exact, int *k,
                       double fidbound, int rank, int fidelity, int forceL, int verbose, int quiet);

/*
 * Weights of the row space of L for decompose's fidelity Z_L
 * (probability.c:371-410): hist[w] = #{i in [0, 2^k) : |sum_j i_j L_j| = w},
 * w = 0..64 (65 entries), on the device.  k <= 40.
 */
int cse_lspace_weights(cse_ctx *ctx, const uint64_t *L, int k, int64_t *hist);

/* exact value of a histogram as FP64 (double-double internally) */
void cse_hist_value(const int64_t *hist, int pmin, double *re, double *im);

/*
 * The sampling-disable rule of probability.c:162-174: returns 1 when the
 * projector norms are to be sampled, 0 when they are computed exactly
 * (noapprox, or samples * bins * 2 > chi - 1 without forceSample).
 */
int cse_sampling_kept(int noapprox, int forceSample, int exact, int t, int k, int samples, int bins);

/*
 * ||P |H^t>||^2 as the backend computes it for G' and H'
 * (probability.c:185-205): exactProjector (innerprod.c:148-199) when
 * sampled == 0, else multiSampledProjector (innerprod.c:23-84) -- the median
 * over bins of the mean over samples, bin b drawing the random states
 * [b * samples, (b + 1) * samples) of the Philox stream keyed by seed.  The
 * empty projector gives norm^2 and t = 0 the Clifford closed form, without
 * the GPU.  *ctx is created on device when first needed (CSE_ERR_NODEV if
 * that fails); the caller destroys it.
 */
int cse_projector_norm(cse_ctx **ctx, int device, const cse_projector *P, const uint64_t *L, int k, int exact,
                       double norm, int sampled, int samples, int bins, uint64_t seed, double *value);

/*
 * cse_projector_norm for nproj projectors of one decomposition (G' and H' of
 * probability.c:185-205, whose two exactProjector / multiSampledProjector
 * calls it replaces): values[i] is cse_projector_norm of P[i] with seed
 * seeds[i].  With sampled == 0 every projector the closed forms leave is one
 * instance of a single cse_batch_exact launch (the values are exact sums, so
 * they equal the single calls' bit for bit); sampled projectors run one
 * after another.  seeds may be NULL when sampled == 0.
 */
int cse_projector_norms(cse_ctx **ctx, int device, int nproj, const cse_projector *P, const uint64_t *seeds,
                        const uint64_t *L, int k, int exact, double norm, int sampled, int samples, int bins,
                        double *values);

#ifdef __cplusplus
}
#endif
#endif
