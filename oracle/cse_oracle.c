/*
 * cse_oracle.c -- CPU oracle (test infrastructure only; see cse_oracle.h).
 *
 * Each function restates the reference algorithm it names (file:line of
 * patrickrall/CircuitSimulator), on uint64 rows instead of the reference's
 * byte-packed MSB-first BitMatrix.  Semantics, including which parts of the
 * n x n J matrix and the D vector get rewritten, follow the reference so that
 * the full-state fixtures of its C unit tests compare equal.
 */
#include "cse_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define BIT(x, i) ((int)(((x) >> (i)) & 1ULL))
#define ONE(i) (1ULL << (i))

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }
static inline int par(uint64_t x) { return __builtin_parityll(x); }
static inline uint64_t lowmask(int k) { return k >= 64 ? ~0ULL : (ONE(k) - 1ULL); }

/* ------------------------------------------------------------------ */
/* stabilizer.c:9-25 -- G = Gbar = I, h = 0, Q = 0, D = 0, J = 0       */
void orc_state_init(orc_state *s, int n, int k)
{
    memset(s, 0, sizeof(*s));
    s->n = n;
    s->k = k;
    for (int i = 0; i < n; i++) {
        s->G[i] = ONE(i);
        s->Gb[i] = ONE(i);
    }
}

/* stabilizer.c:55-62 */
int orc_getD(const orc_state *s, int a) { return 2 * BIT(s->D1, a) + 4 * BIT(s->D2, a); }

void orc_setD(orc_state *s, int a, int val)
{
    int v = ((val % 8) + 8) % 8;
    s->D1 = (s->D1 & ~ONE(a)) | ((uint64_t)((v >> 1) & 1) << a);
    s->D2 = (s->D2 & ~ONE(a)) | ((uint64_t)((v >> 2) & 1) << a);
}

/* stabilizer.c:129-160.  eq. 49: D'_i = sum_{j<k} D_j R_ij, plus
 * 4*[sum_{b>c} J_bc R_ab R_ac odd]; eq. 50: J' = R J R^T over all n rows. */
void orc_updateDJ(orc_state *s, const uint64_t *R)
{
    const int n = s->n;
    const uint64_t km = lowmask(s->k);
    uint64_t D1n = 0, D2n = 0;
    for (int i = 0; i < n; i++) {
        int v = 2 * popc(R[i] & km & s->D1) + 4 * popc(R[i] & km & s->D2);
        v &= 7;
        D1n |= (uint64_t)((v >> 1) & 1) << i;
        D2n |= (uint64_t)((v >> 2) & 1) << i;
    }
    for (int a = 0; a < n; a++) {
        uint64_t T = R[a];
        int cnt = 0;
        for (uint64_t t = T; t; t &= t - 1) {
            int b = __builtin_ctzll(t);
            cnt += popc(s->J[b] & T & lowmask(b));
        }
        if (cnt & 1) D2n ^= ONE(a);
    }
    s->D1 = D1n;
    s->D2 = D2n;

    uint64_t RJ[ORC_NMAX], Jn[ORC_NMAX];
    for (int i = 0; i < n; i++) {
        uint64_t acc = 0;
        for (uint64_t t = R[i]; t; t &= t - 1) acc ^= s->J[__builtin_ctzll(t)];
        RJ[i] = acc;
    }
    for (int i = 0; i < n; i++) {
        uint64_t row = 0;
        for (int j = 0; j < n; j++)
            if (par(RJ[i] & R[j])) row |= ONE(j);
        Jn[i] = row;
    }
    memcpy(s->J, Jn, sizeof(uint64_t) * (size_t)n);
}

/* stabilizer.c:163-177.  eq. 52: Q += sum_{a<k} D_a y_a + 4 sum_{a<b<k} J_ab y_a y_b;
 * eq. 53: D2 ^= J y (all n rows). */
void orc_updateQD(orc_state *s, uint64_t y)
{
    const int k = s->k;
    int Q = s->Q;
    for (int a = 0; a < k; a++) {
        if (!BIT(y, a)) continue;
        Q += orc_getD(s, a);
        for (int b = a + 1; b < k; b++) Q += 4 * (BIT(s->J[a], b) & BIT(y, b));
    }
    s->Q = Q % 8;
    uint64_t t = 0;
    for (int a = 0; a < s->n; a++)
        if (par(s->J[a] & y)) t |= ONE(a);
    s->D2 ^= t;
}

/* stabilizer.c:183-225 -- 1 + w^A + w^B - w^{A+B} for even A, B >= 0 */
void orc_Gamma(int *eps, int *p, int *m, int A, int B)
{
    static const int re[4] = {1, 0, -1, 0};
    static const int im[4] = {0, 1, 0, -1};
    int gr = 1 + re[(A % 8) / 2] + re[(B % 8) / 2] - re[((A + B) % 8) / 2];
    int gi = 0 + im[(A % 8) / 2] + im[(B % 8) / 2] - im[((A + B) % 8) / 2];
    if (gr == 0 && gi == 0) {
        *eps = 0; *p = 0; *m = 0;
        return;
    }
    *eps = 1;
    *p = 2;
    if (gr == 0) {
        *m = gi > 0 ? 2 : 6;
    } else if (gr / 2 == -1) {
        *m = 4;
    } else {
        if (gi / 2 == 1) *m = 1;
        if (gi / 2 == 0) *m = 0;
        if (gi / 2 == -1) *m = 7;
    }
}

/* stabilizer.c:228-255 -- 1 + w^A for even A */
void orc_partialGamma(int *eps, int *p, int *m, int A)
{
    while (A < 0) A += 8;
    switch (A % 8) {
    case 0: *eps = 1; *p = 2; *m = 0; break;
    case 2: *eps = 1; *p = 1; *m = 1; break;
    case 4: *eps = 0; *p = 0; *m = 0; break;
    default: *eps = 1; *p = 1; *m = 7; break; /* 6 */
    }
}

/* stabilizer.c:258-297 */
static void Wsigma(const orc_state *s, int *eps, int *p, int *m, int sigma, int sidx,
                   const int *M, int nM, const int *Dm, int nD)
{
    if (s->k == 0) {
        *eps = 1; *p = 0; *m = s->Q;
        return;
    }
    int tp = 0;
    int tm = s->Q + sigma * orc_getD(s, sidx);
    for (int i = 0; i < nM; i++) {
        int e, pp, mm;
        orc_partialGamma(&e, &pp, &mm, orc_getD(s, M[i]) + sigma * 4 * BIT(s->J[M[i]], sidx));
        if (!e) { *eps = 0; *p = 0; *m = 0; return; }
        tp += pp;
        tm = (tm + mm) % 8;
    }
    for (int i = 0; i < nD; i++) {
        int a = Dm[2 * i], b = Dm[2 * i + 1];
        int e, pp, mm;
        orc_Gamma(&e, &pp, &mm, orc_getD(s, a) + sigma * 4 * BIT(s->J[a], sidx),
                  orc_getD(s, b) + sigma * 4 * BIT(s->J[b], sidx));
        if (!e) { *eps = 0; *p = 0; *m = 0; return; }
        tp += pp;
        tm = (tm + mm) % 8;
    }
    *eps = 1; *p = tp; *m = tm;
}

static void identity_rows(uint64_t *R, int n)
{
    for (int i = 0; i < n; i++) R[i] = ONE(i);
}

/* stabilizer.c:300-481 -- sum_{x in F2^k} w^{q(x)} = eps 2^{p/2} w^m */
void orc_exponentialSum(orc_state *s, int *eps, int *p, int *m)
{
    const int n = s->n, k = s->k;
    int S[ORC_NMAX], nS = 0;
    for (int a = 0; a < k; a++) {
        int d = orc_getD(s, a);
        if (d == 2 || d == 6) S[nS++] = a;
    }
    uint64_t R[ORC_NMAX];
    if (nS > 0) {
        int a = S[0];
        identity_rows(R, n);
        for (int i = 1; i < nS; i++) R[S[i]] ^= ONE(a);
        orc_updateDJ(s, R);
        identity_rows(R, n);
        /* swap columns a and k-1 of the identity */
        R[a] = ONE(k - 1);
        R[k - 1] = ONE(a);
        orc_updateDJ(s, R);
        S[0] = k - 1;
        nS = 1;
    }
    int E[ORC_NMAX], nE = 0;
    for (int x = 0; x < k; x++)
        if (nS == 0 || x != S[0]) E[nE++] = x;

    int M[ORC_NMAX], nM = 0;
    int Dm[2 * ORC_NMAX], nD = 0;
    while (nE > 0) {
        int a = E[0];
        int b = -1;
        for (int i = 1; i < nE; i++)
            if (BIT(s->J[a], E[i])) { b = E[i]; break; }
        if (b < 0) {
            M[nM++] = a;
            for (int i = 0; i < nE - 1; i++) E[i] = E[i + 1];
            nE--;
        } else {
            identity_rows(R, n);
            for (int i = 0; i < nE; i++) {
                int e = E[i];
                if (e == a || e == b) continue;
                if (BIT(s->J[a], e)) R[e] ^= ONE(b);
                if (BIT(s->J[b], e)) R[e] ^= ONE(a);
            }
            orc_updateDJ(s, R);
            Dm[2 * nD] = a;
            Dm[2 * nD + 1] = b;
            nD++;
            int ne = 0;
            for (int i = 0; i < nE; i++)
                if (E[i] != a && E[i] != b) E[ne++] = E[i];
            nE = ne;
        }
    }

    if (nS == 0) {
        Wsigma(s, eps, p, m, 0, 0, M, nM, Dm, nD);
        return;
    }
    int e0, p0, m0, e1, p1, m1;
    Wsigma(s, &e0, &p0, &m0, 0, S[0], M, nM, Dm, nD);
    Wsigma(s, &e1, &p1, &m1, 1, S[0], M, nM, Dm, nD);
    if (e0 == 0) { *eps = e1; *p = p1; *m = m1; return; }
    if (e1 == 0) { *eps = e0; *p = p0; *m = m0; return; }
    /* p0 == p1 and m1 - m0 even always hold here (the reference only prints) */
    orc_partialGamma(eps, p, m, m1 - m0);
    if (*eps == 0) {
        *p = 0;
        *m = 0;
    } else {
        *p += p0;
        *m = *m + m0 % 8;
    }
}

/* stabilizer.c:500-585.  Returns 0 EMPTY, 1 SAME, 2 SUCCESS.  The basis
 * changes R = I + e_a e_i^T and the swap (i, k-1) are applied through
 * orc_updateDJ exactly as the reference does. */
int orc_shrink(orc_state *s, uint64_t xi, int alpha, int lazy)
{
    const int n = s->n;
    int S[ORC_NMAX], nS = 0;
    for (int a = 0; a < s->k; a++)
        if (par(s->G[a] & xi)) S[nS++] = a;
    int beta = (alpha + popc(xi & s->h)) % 2;
    if (nS == 0) return 1 - beta;

    int i = S[--nS];
    uint64_t R[ORC_NMAX];
    uint64_t gbi = s->Gb[i];
    for (int t = 0; t < nS; t++) {
        int a = S[t];
        s->G[a] ^= s->G[i];
        if (lazy != 1) {
            identity_rows(R, n);
            R[a] |= ONE(i);
            orc_updateDJ(s, R);
        }
        gbi ^= s->Gb[a];
    }
    s->Gb[i] = gbi;

    int last = s->k - 1;
    uint64_t tmp = s->G[i]; s->G[i] = s->G[last]; s->G[last] = tmp;
    tmp = s->Gb[i]; s->Gb[i] = s->Gb[last]; s->Gb[last] = tmp;
    if (lazy != 1) {
        identity_rows(R, n);
        R[i] = ONE(last);
        R[last] = ONE(i);
        orc_updateDJ(s, R);
    }
    if (beta) s->h ^= s->G[last];
    if (lazy != 1) orc_updateQD(s, (uint64_t)beta << last);
    s->k--;
    return 2;
}

/* stabilizer.c:589-659.  Result = <s2|s1> = eps 2^{p/2} w^m. */
void orc_innerProduct(const orc_state *s1, const orc_state *s2, int *eps, int *p, int *m)
{
    const int n = s1->n;
    orc_state st = *s1;
    for (int b = s2->k; b < s2->n; b++) {
        uint64_t xi = s2->Gb[b];
        int alpha = par(s2->h & xi);
        *eps = orc_shrink(&st, xi, alpha, 0);
        if (*eps == 0) {
            *p = 0;
            *m = 0;
            return;
        }
    }
    uint64_t R[ORC_NMAX];
    uint64_t hh = st.h ^ s2->h;
    uint64_t y = 0;
    for (int a = 0; a < n; a++) {
        if (par(hh & s2->Gb[a])) y |= ONE(a);
    }
    for (int b = 0; b < n; b++) {
        uint64_t row = 0;
        for (int a = 0; a < n; a++)
            if (par(st.G[b] & s2->Gb[a])) row |= ONE(a);
        R[b] = row;
    }
    orc_state t2 = *s2;
    orc_updateQD(&t2, y);
    orc_updateDJ(&t2, R);

    st.Q -= t2.Q;
    if (st.Q < 0) st.Q += 8;
    for (int i = 0; i < n; i++) {
        int v = orc_getD(&st, i) - orc_getD(&t2, i);
        if (v < 0) v += 8;
        orc_setD(&st, i, v);
    }
    for (int i = 0; i < n; i++) st.J[i] ^= t2.J[i];

    orc_exponentialSum(&st, eps, p, m);
    *p -= s1->k + s2->k;
}

/* stabilizer.c:759-825 */
void orc_extend(orc_state *s, uint64_t xi)
{
    const int n = s->n;
    int S[ORC_NMAX], nS = 0;
    for (int a = 0; a < n; a++)
        if (par(xi & s->Gb[a])) S[nS++] = a;
    int i = -1;
    for (int t = 0; t < nS; t++)
        if (s->k <= S[t] && s->k < n) { i = S[t]; break; }
    if (i < 0) return;
    uint64_t gbi = s->Gb[i], gi = s->G[i];
    for (int t = 0; t < nS; t++) {
        int a = S[t];
        if (a == i) continue;
        s->Gb[a] ^= gbi;
        gi ^= s->G[a];
    }
    s->G[i] = gi;
    int kk = s->k;
    uint64_t tmp = s->G[i]; s->G[i] = s->G[kk]; s->G[kk] = tmp;
    tmp = s->Gb[i]; s->Gb[i] = s->Gb[kk]; s->Gb[kk] = tmp;
    s->k++;
}

/* stabilizer.c:827-959; P = i^m Z(zeta) X(xi) */
int orc_measurePauli(orc_state *s, int m, uint64_t zeta, uint64_t xi)
{
    const int n = s->n, k = s->k;
    uint64_t vxi = 0, vze = 0;
    for (int a = 0; a < k; a++) {
        if (par(s->Gb[a] & xi)) vxi |= ONE(a);
        if (par(s->G[a] & zeta)) vze |= ONE(a);
    }
    uint64_t xip = 0;
    for (int a = 0; a < k; a++)
        if (BIT(vxi, a)) xip ^= s->G[a];

    int w = 2 * m + 4 * par(zeta & s->h);
    for (int b = 0; b < k; b++) w += orc_getD(s, b) * BIT(vxi, b);
    for (int b = 0; b < k; b++)
        for (int a = 0; a < b; a++) w += 4 * (BIT(s->J[a], b) & BIT(vxi, a) & BIT(vxi, b));
    w %= 8;

    uint64_t eta = 0;
    for (int a = 0; a < n; a++)
        if (par(s->J[a] & vxi)) eta |= ONE(a);
    eta ^= vze;

    if (xi == xip) {
        if (w == 0 || w == 4) {
            uint64_t gamma = 0;
            for (int a = 0; a < k; a++)
                if (BIT(eta, a)) gamma ^= s->Gb[a];
            int alpha = (w / 4 + popc(gamma & s->h)) % 2;
            return orc_shrink(s, gamma, alpha, 0);
        } else {
            int sigma = 2 - w / 2;
            s->Q = (s->Q + sigma) % 8;
            while (s->Q < 0) s->Q += 8;
            for (int a = 0; a < k; a++) {
                int v = orc_getD(s, a) - 2 * sigma * BIT(eta, a);
                while (v < 0) v += 8;
                orc_setD(s, a, v);
            }
            for (int i = 0; i < n; i++)
                if (BIT(eta, i)) s->J[i] ^= eta;
            return 2;
        }
    }
    orc_extend(s, xi);
    int newD = 2 * m + 4 * par(zeta & xi) + 4 * par(zeta & s->h);
    int last = s->k - 1;
    orc_setD(s, last, newD);
    /* J row last := vzeta, column last := vzeta, J_last,last := m mod 2 */
    s->J[last] = vze & lowmask(n);
    for (int i = 0; i < n; i++) {
        if (BIT(vze, i)) s->J[i] |= ONE(last);
        else s->J[i] &= ~ONE(last);
    }
    if (m % 2) s->J[last] |= ONE(last);
    else s->J[last] &= ~ONE(last);
    return 2;
}

/* ------------------------------------------------------------------ */
/* stateprep.c:36-81.  size = ceil(t/2); bit j of the MSB-first binary
 * representation of i (binrep, stateprep.c:5-32) selects pair j. */
void orc_prepH(orc_state *s, uint64_t i, int t)
{
    int size = (t + 1) / 2;
    orc_state_init(s, t, t);
    for (int j = 0; j < size; j++) {
        int bit = (int)((i >> (size - 1 - j)) & 1ULL);
        if (bit == 0 && !((t % 2) && j == size - 1)) {
            s->J[2 * j + 1] |= ONE(2 * j);
            s->J[2 * j] |= ONE(2 * j + 1);
        }
    }
    for (int j = 0; j < size; j++) {
        int bit = (int)((i >> (size - 1 - j)) & 1ULL);
        if ((t % 2) && j == size - 1) {
            if (bit) orc_shrink(s, ONE(t - 1), 0, 0);
            continue;
        }
        if (bit) orc_shrink(s, ONE(2 * j) | ONE(2 * j + 1), 0, 0);
    }
}

/* stateprep.c:85-120.  x = sum_j bit_{k-1-j}(i) L[j]; |0> where x_q = 0. */
void orc_prepL(orc_state *s, uint64_t i, int t, const uint64_t *L, int k)
{
    uint64_t x = 0;
    for (int j = 0; j < k; j++)
        if ((i >> (k - 1 - j)) & 1ULL) x ^= L[j];
    orc_state_init(s, t, t);
    for (int q = 0; q < t; q++)
        if (!BIT(x, q)) orc_shrink(s, ONE(q), 0, 0);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10, from Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl
 * key increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* word w of stream: block (w>>2) at counter {w>>2, stream lo, stream hi, domain} */
uint32_t orc_rand32(const uint32_t key[2], uint64_t stream, uint32_t domain, uint64_t w)
{
    uint32_t ctr[4] = {(uint32_t)(w >> 2), (uint32_t)stream, (uint32_t)(stream >> 32), domain};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[w & 3];
}

/* stabilizer.c:677-716: eta(d) ~ 2^{-d(d+1)/2} prod_{a=1..d} (1-2^{d-n-a})/(1-2^{-a}),
 * normalised, cumulated, normalised again; thresholds scaled by 2^53. */
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

void orc_dimension_thresholds(int n, uint64_t *thr)
{
    double dist[ORC_NMAX + 1], cum[ORC_NMAX + 1];
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

#define ORC_RS_CAP 4096
#define ORC_DOM_STATE 1u

/* stabilizer.c:689-756 with the RNG replaced by a Philox stream:
 *   word 0,1   -> x (53 bits); d = min{d : x + 1 <= thr[d]}; k = n - d
 *   then lazy shrinks by random xi (2 words each) until dim k (cap 4096)
 *   then h, D1, D2 (2 words each), then J lower rows i < k (2 words each);
 *   J_ii = [D_i in {2,6}], Q = 0 (the C reference leaves Q at 0). */
int orc_random_state(orc_state *s, int n, const uint64_t *thr, const uint32_t key[2],
                     uint64_t stream)
{
    uint64_t w = 0;
#define R64() (w += 2, ((uint64_t)orc_rand32(key, stream, ORC_DOM_STATE, w - 2) | \
                        ((uint64_t)orc_rand32(key, stream, ORC_DOM_STATE, w - 1) << 32)))
    const uint64_t nm = lowmask(n);
    uint64_t x = R64() >> 11;
    int d = 0;
    while (d < n && !(x + 1 <= thr[d])) d++;
    int k = n - d;
    orc_state_init(s, n, n);
    int iters = 0;
    while (s->k > k) {
        if (++iters > ORC_RS_CAP) return -1;
        uint64_t xi = R64() & nm;
        orc_shrink(s, xi, 0, 1);
    }
    s->h = R64() & nm;
    s->D1 = R64() & nm;
    s->D2 = R64() & nm;
    for (int i = 0; i < k; i++) {
        uint64_t lo = R64() & lowmask(i);
        s->J[i] |= lo;
        for (uint64_t t = lo; t; t &= t - 1) s->J[__builtin_ctzll(t)] |= ONE(i);
        if (BIT(s->D1, i)) s->J[i] |= ONE(i);
    }
    s->Q = 0;
#undef R64
    return (int)w;
}

/* ------------------------------------------------------------------ */
/* innerprod.c:98-116: projfactor = prod of measurePauli results */
int orc_project(orc_state *s, const orc_projector *P)
{
    int sexp = 0;
    for (int g = 0; g < P->nstabs; g++) {
        int r = orc_measurePauli(s, P->phase[g] & 3, P->zs[g], P->xs[g]);
        if (r == 0) return -1;
        if (r == 2) sexp++;
    }
    return sexp;
}

int orc_exact_term(uint64_t l, const orc_projector *P, const uint64_t *L, int k, int exact,
                   uint64_t *i_out, uint64_t *j_out, orc_ipres *ip, int *s_out)
{
    const int t = P->nqubits;
    uint64_t chi = exact ? (1ULL << ((t + 1) / 2)) : (1ULL << k);
    uint64_t i = 0;
    while (l >= chi - i) {
        l -= chi - i;
        i++;
    }
    uint64_t j = l + i;
    *i_out = i;
    *j_out = j;
    orc_state th, ph;
    if (exact) orc_prepH(&th, i, t);
    else orc_prepL(&th, i, t, L, k);
    int sexp = orc_project(&th, P);
    if (sexp < 0) {
        ip->eps = 0; ip->p = 0; ip->m = 0;
        *s_out = 0;
        return 1;
    }
    if (exact) orc_prepH(&ph, j, t);
    else orc_prepL(&ph, j, t, L, k);
    orc_innerProduct(&th, &ph, &ip->eps, &ip->p, &ip->m);
    *s_out = sexp;
    return 0;
}

/* ------------------------------------------------------------------ */
void orc_hist_add(int64_t *hist, int P, int M, int64_t count)
{
    int mm = ((M % 8) + 8) % 8;
    hist[(-P) * 8 + mm] += count;
}

/* 2^{p/2} w^m = 2^{floor(p/2)} (sqrt2)^{p&1} w^m, sqrt2 = w - w^3, w^4 = -1.
 * Adds eps * 2^{E + floor(p/2)} times the Z[w] coordinates to c. */
void orc_zomega_add(int64_t c[4], int E, int eps, int p, int m)
{
    if (!eps) return;
    int e = E + (p >> 1); /* arithmetic shift = floor(p/2) */
    int64_t v = (int64_t)1 << e;
    int mm = ((m % 8) + 8) % 8;
    int terms[2], nt = 0;
    int signs[2];
    if ((p & 1) == 0) {
        terms[nt] = mm; signs[nt++] = 1;
    } else {
        terms[nt] = mm + 1; signs[nt++] = 1;
        terms[nt] = mm + 3; signs[nt++] = -1;
    }
    for (int u = 0; u < nt; u++) {
        int q = terms[u] % 8;
        int sg = signs[u];
        if (q >= 4) { q -= 4; sg = -sg; }
        c[q] += sg * v;
    }
}

void orc_zomega_norm2(const int64_t c[4], __int128 *A, __int128 *B)
{
    __int128 c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    *A = c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3;
    *B = c0 * c1 - c0 * c3 + c1 * c2 + c2 * c3;
}

/* ------------------------------------------------------------------ */
/* Dense-row cost model: a shrink step on a k-dimensional state costs
 * 9 word ops per active row (ballot term, row XOR, J row and column update,
 * D update); the basis alignment costs 6 word ops per (row, column) of the
 * k x k2 map; the exponential sum costs 8 word ops per remaining row per
 * elimination step.  Counted while running the same shrink sequence. */
long orc_ip_wordops(const orc_state *s1, const orc_state *s2)
{
    orc_state st = *s1;
    long ops = 0;
    for (int b = s2->k; b < s2->n; b++) {
        ops += 9L * st.k;
        uint64_t xi = s2->Gb[b];
        int e = orc_shrink(&st, xi, par(s2->h & xi), 0);
        if (e == 0) return ops;
    }
    ops += 6L * st.k * s2->k;
    ops += 8L * st.k * st.k;
    return ops;
}

/* ------------------------------------------------------------------ */
/* Sampled-mode sweep for given projected thetas (innerprod.c:119-139):
 * per theta, sum_{i < nphi} <prepL(i) | theta'> as Z[w] fixed-point sums at
 * scale 2^E, and optionally every term packed as 0 (eps = 0) or
 * (-p) * 8 + (m mod 8) + 1 at terms[theta * nphi + i].  nthreads POSIX
 * threads split the thetas cyclically (the rank-strided loop of
 * probability.c:268).  Returns the number of inner products. */
typedef struct {
    const orc_state *thetas;
    int ntheta, t, k, E, nphi, nthreads, tid;
    const uint64_t *L;
    int64_t *csum;
    int32_t *terms;
    long count;
} orc_sweep_job;

static void *orc_sweep_worker(void *arg)
{
    orc_sweep_job *j = (orc_sweep_job *)arg;
    for (int th = j->tid; th < j->ntheta; th += j->nthreads) {
        int64_t c[4] = {0, 0, 0, 0};
        for (int i = 0; i < j->nphi; i++) {
            orc_state phi;
            orc_prepL(&phi, (uint64_t)i, j->t, j->L, j->k);
            int eps, p, m;
            orc_innerProduct(&j->thetas[th], &phi, &eps, &p, &m);
            orc_zomega_add(c, j->E, eps, p, m);
            if (j->terms)
                j->terms[(long)th * j->nphi + i] = eps ? (-p) * 8 + (((m % 8) + 8) % 8) + 1 : 0;
            j->count++;
        }
        memcpy(&j->csum[4 * th], c, sizeof(c));
    }
    return NULL;
}

long orc_sampled_sweep(const orc_state *thetas, int ntheta, int t, const uint64_t *L, int k, int E,
                       int nthreads, int nphi, int64_t *csum, int32_t *terms)
{
    if (nphi > (1 << k) || nphi < 0) nphi = 1 << k;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    orc_sweep_job jobs[256];
    for (int i = 0; i < nthreads; i++) {
        jobs[i] = (orc_sweep_job){thetas, ntheta, t, k, E, nphi, nthreads, i, L, csum, terms, 0};
        pthread_create(&th[i], NULL, orc_sweep_worker, &jobs[i]);
    }
    long total = 0;
    for (int i = 0; i < nthreads; i++) {
        pthread_join(th[i], NULL);
        total += jobs[i].count;
    }
    return total;
}
