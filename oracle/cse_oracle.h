/*
 * cse_oracle.h -- CPU oracle for the stabilizer-rank inner-product engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch CPU restatement of the
 * reference algorithm (libcirc/stabilizer/stabilizer.c, libcirc/stateprep.c,
 * libcirc/innerprod.c, libcirc/probability.c of patrickrall/CircuitSimulator)
 * on uint64 rows.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product path
 * (circuitsimulator_amd/) never links, calls or falls back to it.
 *
 * Representation (semantics of stabilizer.h:5-16 and stabilizer.py:54-75):
 *   |K,q> = 2^{-k/2} sum_{x in K} omega^{q(x)} |x>,  omega = e^{i pi/4}
 *   K = h + span(G[0..k-1]),  Gbar = (G^{-1})^T,
 *   q(h + sum_a z_a G[a]) = Q + sum_a D_a z_a + 4 sum_{a<b} J_ab z_a z_b (mod 8)
 * Coordinate j of an n-bit vector is bit j of a uint64 (bit 0 = x[0]).  The
 * brute-force checks index amplitudes by that uint64 (x[0] least significant,
 * as the reference's Unpack fixtures do; stabilizer.py:54-75 reverses the bit
 * order, a qubit relabelling that leaves every inner product unchanged).
 * D_a = 2*D1_a + 4*D2_a (stabilizer.c:55-62).  J is stored as bits (J_ab in
 * {0,4} is bit b of J[a]); the diagonal follows J_aa = [D_a in {2,6}].
 *
 * Pinning (how each function is checked):
 *   exponentialSum  -- exponentialSumTests.txt (274 records) + brute force
 *   shrink          -- shrinkTests.txt (100 records, full n x n state)
 *   measurePauli    -- measurePauliTests.txt (30) + brute force projector
 *   extend          -- extendTests.txt (10)
 *   updateQD        -- QDtests.json (99)
 *   innerProduct    -- innerProductTests-backup.txt (1) + brute force n<=12
 *                      + the compiled reference C (oracle/_ref) on seeded pairs
 *   prepH / prepL   -- closed forms (sum_i phi_i ~ |H>^t, product |0>/|+>)
 *   random state    -- invariants (G Gbar^T = I, J symmetric, J_aa rule,
 *                      dimension law vs eta(d), E|<theta|psi>|^2 = 2^-n)
 *   philox4x32-10   -- Random123 known-answer vectors
 *   projector sums  -- brute-force ||Pi |H^t>||^2 for small t
 */
#ifndef CSE_ORACLE_H
#define CSE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMAX 64

typedef struct {
    int n, k, Q;
    uint64_t h, D1, D2;
    uint64_t G[ORC_NMAX], Gb[ORC_NMAX], J[ORC_NMAX];
} orc_state;

typedef struct {
    int eps, p, m;
} orc_ipres;

/* ---- primitives (stabilizer.c) ---- */
void orc_state_init(orc_state *s, int n, int k);               /* stabilizer.c:9-25 */
int  orc_getD(const orc_state *s, int a);                      /* stabilizer.c:55-57 */
void orc_setD(orc_state *s, int a, int val);                   /* stabilizer.c:59-62 */
void orc_updateDJ(orc_state *s, const uint64_t *R);            /* stabilizer.c:129-160 */
void orc_updateQD(orc_state *s, uint64_t y);                   /* stabilizer.c:163-177 */
void orc_Gamma(int *eps, int *p, int *m, int A, int B);        /* stabilizer.c:183-225 */
void orc_partialGamma(int *eps, int *p, int *m, int A);        /* stabilizer.c:228-255 */
void orc_exponentialSum(orc_state *s, int *eps, int *p, int *m); /* stabilizer.c:300-481 */
int  orc_shrink(orc_state *s, uint64_t xi, int alpha, int lazy); /* stabilizer.c:500-585 */
void orc_innerProduct(const orc_state *s1, const orc_state *s2,
                      int *eps, int *p, int *m);               /* stabilizer.c:589-659 */
void orc_extend(orc_state *s, uint64_t xi);                    /* stabilizer.c:759-825 */
/* returns 0 (norm 0), 1 (norm 1) or 2 (norm 2^-1/2); stabilizer.c:827-959 */
int  orc_measurePauli(orc_state *s, int m, uint64_t zeta, uint64_t xi);

/* ---- state preparation (stateprep.c) ---- */
void orc_prepH(orc_state *s, uint64_t i, int t);               /* stateprep.c:36-81 */
void orc_prepL(orc_state *s, uint64_t i, int t,
               const uint64_t *L, int k);                      /* stateprep.c:85-120 */

/* ---- counter-based RNG: Philox4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* word w of the stream (key, stream id, domain) */
uint32_t orc_rand32(const uint32_t key[2], uint64_t stream, uint32_t domain, uint64_t w);

/* ---- random stabilizer state (stabilizer.c:669-756, Philox-driven) ---- */
/* thr[d] = floor(cdf[d] * 2^53) for d = 0..n (host CDF of eq. 79) */
void orc_dimension_thresholds(int n, uint64_t *thr);
/* returns number of RNG words consumed, or -1 if the lazy-shrink cap was hit */
int  orc_random_state(orc_state *s, int n, const uint64_t *thr,
                      const uint32_t key[2], uint64_t stream);

/* ---- projector (comms.h:4-11): generator g is i^m Z(zs[g]) X(xs[g]) ---- */
typedef struct {
    int nstabs, nqubits;
    int phase[ORC_NMAX];          /* m = 2*phaseSign + phaseComplex (innerprod.c:101-102) */
    uint64_t xs[ORC_NMAX], zs[ORC_NMAX];
} orc_projector;

/* Projects s onto the projector; returns s (the count of 2^-1/2 factors) or -1
 * if the projection vanishes (innerprod.c:98-116 / :222-241). */
int  orc_project(orc_state *s, const orc_projector *P);

/* One exact-mode term l (innerprod.c:203-261).  Out: i, j, the inner product
 * <phi_j | theta_i'> as (eps,p,m) and the projection exponent s
 * (projfactor = 2^{-s/2}); returns 0, or 1 if the projection vanished. */
int  orc_exact_term(uint64_t l, const orc_projector *P, const uint64_t *L, int k,
                    int exact, uint64_t *i_out, uint64_t *j_out,
                    orc_ipres *ip, int *s_out);

/* ---- exact accumulation ----
 * A value sum c * 2^{P/2} omega^M is held as int64 counts per bin
 * (P in [pmin, 0], M in 0..7): bin index (-P)*8 + M. */
void orc_hist_add(int64_t *hist, int P, int M, int64_t count);

/* Z[omega] fixed point: value = 2^{-E} (c0 + c1 w + c2 w^2 + c3 w^3). */
void orc_zomega_add(int64_t c[4], int E, int eps, int p, int m);

/* Exact sampled estimator for one theta: the per-theta |sum|^2 as
 * (A + B sqrt2) 2^{-2E}; computed in __int128. */
void orc_zomega_norm2(const int64_t c[4], __int128 *A, __int128 *B);

/* ---- op-count model of the canonical dense-row algorithm ---- */
/* Counts 64-bit word operations of one inner product of (s1, s2) under a
 * dense-row cost model of innerProductExact (stabilizer.c:589-659): 9 k1 per
 * shrink by a row of s2's dual basis (until one empties the state), 6 k1 k2
 * for the basis alignment and 8 k1^2 for the exponential sum, with k1 the
 * dimension after the shrinks. */
long orc_ip_wordops(const orc_state *s1, const orc_state *s2);

/* ---- sampled-mode sweep over given projected thetas (innerprod.c:119-139) ----
 * csum[4 * theta] = sum_{i < nphi} <prepL(i) | theta'> at Z[w] scale 2^E;
 * terms (may be NULL): packed per term as in the engine's per-term output.
 * nthreads POSIX threads; returns the number of inner products. */
long orc_sampled_sweep(const orc_state *thetas, int ntheta, int t, const uint64_t *L, int k, int E,
                       int nthreads, int nphi, int64_t *csum, int32_t *terms);

#ifdef __cplusplus
}
#endif
#endif
