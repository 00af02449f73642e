// cse_kernels.hip -- gfx950 kernels of the stabilizer-rank inner-product
// engine.  See cse_device.h for the execution model.
//
// Kernels
//   k_ip_pairs       general <s2|s1> for packed state pairs, one wave per pair
//   k_prep_states    prepL / prepH states (reference row bookkeeping)
//   k_random_states  Philox random stabilizer states, projected by P
//   k_records        theta' = Pi phi_i or Pi theta_s, turned into an
//                    elimination-ready record (reduced row echelon basis,
//                    rows indexed by pivot column)
//   k_sweep          persistent: a device work queue of (record, chunk) items;
//                    every wave computes <phi_j | theta'> for the chunk's j,
//                    phi_j generated on the fly from (L, j) or (t, j), and
//                    accumulates exactly (int64) plus a Kahan-FP32 shadow.
//   k_sweep_packed   the same with D packed into the J rows (n <= 60); the
//                    |L> terms of records with at most two non-pivot columns
//                    run 64 at a time in one asm block (phiL_batch), every
//                    |H^t> term in phiH_batch
//   k_sweep_packed_gen  the sampled packed sweep with exact mode's general
//                    restriction loop; aligned batches of |H^t> terms run as
//                    eight groups of 8 that share their pair operations
//   k_sweep_packed_all, k_sweep_packed_all8  the packed sweep of exact mode
//                    (seven / eight waves per SIMD): every |L> term in the
//                    asm block, any number of restrictions, every |H^t> term
//                    after its batch's and group's shared pair operations
//   k_zero_list      clears a call's result and queue blocks in one launch
//   k_lspace_hist    histogram of |x| over the row space of L (decompose's
//                    fidelity)
#include "cse_device.h"
#include "cse_internal.h"

using namespace cse;

namespace {

// ------------------------------------------------------ packed state I/O
DEV void load_rstate(RState &s, const uint64_t *w) {
    const int x = lane_id();
    uint64_t hdr = w[0];
    s.n = (int)(hdr & 0xff);
    s.k = (int)((hdr >> 8) & 0xff);
    s.Q = (int)((hdr >> 16) & 7);
    s.h = w[1];
    u64 D1 = w[2], D2 = w[3];
    s.d = 2 * (int)((D1 >> x) & 1ull) + 4 * (int)((D2 >> x) & 1ull);
    const int n = s.n;
    s.G = x < n ? w[4 + x] : 0ull;
    s.Gb = x < n ? w[4 + n + x] : 0ull;
    s.J = x < n ? w[4 + 2 * n + x] : 0ull;
}

DEV void store_rstate(const RState &s, uint64_t *w) {
    const int x = lane_id();
    const int n = s.n;
    u64 D1 = ballot(x < n && (s.d & 2));
    u64 D2 = ballot(x < n && (s.d & 4));
    if (x == 0) {
        w[0] = (uint64_t)n | ((uint64_t)s.k << 8) | ((uint64_t)(s.Q & 7) << 16);
        w[1] = s.h;
        w[2] = D1;
        w[3] = D2;
    }
    if (x < n) {
        w[4 + x] = s.G;
        w[4 + n + x] = s.Gb;
        w[4 + 2 * n + x] = s.J;
    }
}

DEV void to_astate(const RState &r, AState &s) {
    const int x = lane_id();
    s.G = r.G;
    s.J = r.J;
    s.d = x < r.n ? (r.d & 6) : 0;
    s.act = lowmask(r.k);
    s.h = r.h;
    s.Q = r.Q;
}

// General alignment of innerProductExact (stabilizer.c:611-652) in the
// active-mask layout: the quadratic form of s2, rewritten in the basis of
// the restricted s1 (updateQD(y) then updateDJ(R) on a copy of s2), is
// subtracted from s1's form.
DEV void as_subtract_general(AState &s, const RState &s2) {
    const int x = lane_id();
    const int n = s2.n, k2 = s2.k;
    const u64 k2m = lowmask(k2);
    // y_b = (h + h2) . Gbar2_b
    const u64 y = ballot(x < n && par64((s.h ^ s2.h) & s2.Gb));
    // updateQD(y) on the copy of s2 (stabilizer.c:163-177)
    const bool inY = x < k2 && ((y >> x) & 1ull);
    int Q2 = s2.Q;
    Q2 += 2 * __popcll(ballot(inY && (s2.d & 2))) + 4 * __popcll(ballot(inY && (s2.d & 4)));
    Q2 += 4 * (__popcll(ballot(inY && par64(s2.J & y & k2m & ~lowmask(x + 1)))) & 1);
    Q2 &= 7;
    int d2 = s2.d;
    if (x < n && par64(s2.J & y)) d2 ^= 4;
    // R_a = (G_a . Gbar2_b)_b for the active rows a (zero beyond k2)
    u64 R = 0;
    for (int b = 0; b < k2; b++) {
        u64 g = rl64(s2.Gb, b);
        R |= (u64)par64(s.G & g) << b;
    }
    // D2'_a = sum_{b<k2} D2_b R_ab + 4 [sum_{b>c} J2_bc R_ab R_ac odd]
    const u64 M1 = ballot(x < k2 && (d2 & 2));
    const u64 M2 = ballot(x < k2 && (d2 & 4));
    int dp = 2 * __popcll(R & M1) + 4 * __popcll(R & M2);
    u64 RJ = 0;
    int cross = 0;
    for (int b = 0; b < k2; b++) {
        u64 jb = rl64(s2.J, b);
        if ((R >> b) & 1ull) {
            RJ ^= jb;
            cross ^= par64(jb & R & lowmask(b));
        }
    }
    dp = (dp + 4 * cross) & 7;
    // J2' = R J2 R^T on the active rows
    u64 J2p = 0;
    u64 act = s.act;
    while (act) {
        int a2 = __ffsll((long long)act) - 1;
        act &= act - 1;
        u64 Ra2 = rl64(R, a2);
        J2p |= (u64)par64(RJ & Ra2) << a2;
    }
    // D -= D2', J ^= J2', Q -= Q2'
    s.d = (s.d - dp + 8) & 7;
    s.J ^= J2p;
    s.Q = (s.Q - Q2 + 8) & 7;
}

// The restrictions of innerProductExact (stabilizer.c:594-607) on the packed
// layout (D_x in bits 61..63 of J_x, n <= 60), one asm block: for each row b
// of Gbar2 from k2 to n - 1, xi = that row and the right side alpha =
// h2 . xi; the column is the ballot of the parity of G & xi over the active
// rows, c = (h2 ^ h) . xi, and the elimination of i = min(S) over S is the
// exact-mode sweep's (PHIL_RESTRICT_ALL): rows of S take J_i, D_i + 4 J_xi
// and G_i, the rows of M = J_i (^ S when D_i is in {2,6}) take S, and c = 1
// adds D_i to Q (bits 29..31 of s60), 4 to D on M and G_i to h.  Returns 0
// when a restriction is empty with right side 1 (the product is 0).
DEV int pk_restrict_general(unsigned &jlo, unsigned &jhi, unsigned &glo, unsigned &ghi, u64 &h, u64 &act,
                            unsigned &W, unsigned gblo, unsigned gbhi, u64 h2, unsigned bstart, unsigned bend) {
    unsigned vt0, vt1, vt2, vt3;
    int ok;
    asm volatile(
        "s_mov_b32 %[ok], 1\n\t"
        "s_mov_b32 s50, %[bstart]\n"
        "L_gb%=:\n\t"
        "s_cmp_ge_u32 s50, %[bend]\n\t"
        "s_cbranch_scc1 L_gdone%=\n\t"
        "v_readlane_b32 s52, %[gblo], s50\n\t"
        "v_readlane_b32 s53, %[gbhi], s50\n\t"
        "s_add_u32 s50, s50, 1\n\t"
        "s_xor_b64 s[66:67], %[h2], s[48:49]\n\t"
        "s_and_b64 s[66:67], s[66:67], s[52:53]\n\t"
        "s_bcnt1_i32_b64 s51, s[66:67]\n\t"
        "v_and_b32 %[vt0], s52, %[glo]\n\t"
        "v_and_b32 %[vt1], s53, %[ghi]\n\t"
        "v_bcnt_u32_b32 %[vt0], %[vt0], 0\n\t"
        "v_bcnt_u32_b32 %[vt0], %[vt1], %[vt0]\n\t"
        "v_and_b32 %[vt0], 1, %[vt0]\n\t"
        "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t"
        "s_and_b64 s[64:65], s[52:53], s[58:59]\n\t"
        "s_cbranch_scc0 L_gempty%=\n\t"
        "s_ff1_i32_b64 s56, s[64:65]\n\t"
        "s_bitset0_b64 s[58:59], s56\n\t"
        "v_readlane_b32 s66, %[jlo], s56\n\t"
        "v_readlane_b32 s67, %[jhi], s56\n\t"
        "v_readlane_b32 s46, %[glo], s56\n\t"
        "v_readlane_b32 s47, %[ghi], s56\n\t"
        "s_and_b32 s56, s67, 0xe0000000\n\t"
        "s_and_b32 s67, s67, 0x1fffffff\n\t"
        "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t"
        "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t"
        "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t"
        "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t"
        "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t"
        "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t"
        "v_bitop3_b32 %[glo], %[glo], s46, %[vt0] bitop3:0x78\n\t"
        "v_bitop3_b32 %[ghi], %[ghi], s47, %[vt0] bitop3:0x78\n\t"
        "s_bitcmp1_b32 s56, 30\n\t"
        "s_cbranch_scc0 L_gnojii%=\n\t"
        "v_xor_b32 %[vt3], %[vt3], %[vt0]\n"
        "L_gnojii%=:\n\t"
        "s_bitcmp1_b32 s51, 0\n\t"
        "s_cbranch_scc0 L_gnoc%=\n\t"
        "s_or_b32 s65, s65, 0x80000000\n\t"
        "s_add_u32 s60, s60, s56\n\t"
        "s_xor_b64 s[48:49], s[48:49], s[46:47]\n"
        "L_gnoc%=:\n\t"
        "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t"
        "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n\t"
        "s_branch L_gb%=\n"
        "L_gempty%=:\n\t"
        "s_bitcmp1_b32 s51, 0\n\t"
        "s_cbranch_scc0 L_gb%=\n\t"
        "s_mov_b32 %[ok], 0\n"
        "L_gdone%=:\n\t"
        : [ok] "=&s"(ok), [jlo] "+v"(jlo), [jhi] "+v"(jhi), [glo] "+v"(glo), [ghi] "+v"(ghi),
          "+{s[48:49]}"(h), "+{s[58:59]}"(act), "+{s60}"(W), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
          [vt2] "=&v"(vt2), [vt3] "=&v"(vt3)
        : [gblo] "v"(gblo), [gbhi] "v"(gbhi), [h2] "s"(h2), [bstart] "s"(bstart), [bend] "s"(bend)
        : "s46", "s47", "s50", "s51", "s52", "s53", "s56", "s64", "s65", "s66", "s67", "scc");
    return ok;
}

// <s2|s1> for general packed states (stabilizer.c:589-659).  For n <= 60
// the restrictions run in pk_restrict_general and the exponential sum in
// the hand-scheduled loop (as_expsum_packed), on the packed layout.
DEV void ip_general(const RState &r1, const RState &r2, int &eps, int &p, int &m) {
    AState s;
    to_astate(r1, s);
    if (r1.n <= CSE_END_ROW) {
        const int x = lane_id();
        unsigned jlo = (unsigned)s.J, jhi = (unsigned)(s.J >> 32) | ((unsigned)s.d << 29);
        unsigned glo = (unsigned)s.G, ghi = (unsigned)(s.G >> 32);
        u64 h = uniform64(s.h), act = uniform64(s.act);
        unsigned W = (unsigned)__builtin_amdgcn_readfirstlane((int)((unsigned)s.Q << 29));
        const int ok = pk_restrict_general(jlo, jhi, glo, ghi, h, act, W, (unsigned)r2.Gb, (unsigned)(r2.Gb >> 32),
                                           uniform64(r2.h), (unsigned)__builtin_amdgcn_readfirstlane(r2.k),
                                           (unsigned)__builtin_amdgcn_readfirstlane(r2.n));
        if (!ok) {
            eps = 0; p = 0; m = 0;
            return;
        }
        s.J = ((u64)(jhi & 0x1fffffffu) << 32) | jlo;
        s.d = (int)(jhi >> 29);
        s.G = ((u64)ghi << 32) | glo;
        s.h = h;
        s.act = act;
        s.Q = (int)(W >> 29);
        as_subtract_general(s, r2);
        // packed again for the loop: D in bits 61..63, row 60 the end row
        s.J = (s.J & lowmask(CSE_END_ROW)) | ((u64)(s.d & 7) << 61);
        if (x == CSE_END_ROW) s.J = 4ull << 61;
        int pp, mm;
        eps = as_expsum_packed(s, pp, mm);
        p = eps ? pp - r1.k - r2.k : 0;
        m = mm;
        return;
    }
    for (int b = r2.k; b < r2.n; b++) {
        u64 xi = rl64(r2.Gb, b);
        int alpha = par64(r2.h & xi);
        if (!as_restrict<false>(s, xi, alpha)) {
            eps = 0; p = 0; m = 0;
            return;
        }
    }
    as_subtract_general(s, r2);
    int pp, mm;
    eps = as_expsum<false>(s, pp, mm);
    p = eps ? pp - r1.k - r2.k : 0;
    m = mm;
}

}  // namespace

// ======================================================================
extern "C" __global__ void __launch_bounds__(256)
k_ip_pairs(const uint64_t *__restrict__ s1, const uint64_t *__restrict__ s2, int n, int npairs,
           int32_t *__restrict__ out) {
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (wave >= npairs) return;
    const size_t stride = (size_t)(3 * n + 4);
    RState a, b;
    load_rstate(a, s1 + stride * wave);
    load_rstate(b, s2 + stride * wave);
    int eps, p, m;
    ip_general(a, b, eps, p, m);
    if (lane_id() == 0) {
        out[3 * wave] = eps;
        out[3 * wave + 1] = p;
        out[3 * wave + 2] = m;
    }
}

namespace {

// stateprep.c:85-120 in the reference row bookkeeping
DEV void rs_prepL(RState &s, u64 i, int t, const uint64_t *L, int k) {
    u64 xv = 0;
    for (int j = 0; j < k; j++)
        if ((i >> (k - 1 - j)) & 1ull) xv ^= L[j];
    rs_init(s, t);
    for (int q = 0; q < t; q++)
        if (!((xv >> q) & 1ull)) rs_shrink(s, bit(q), 0, false);
}

// stateprep.c:36-81 in the reference row bookkeeping
DEV void rs_prepH(RState &s, u64 i, int t) {
    const int x = lane_id();
    const int size = (t + 1) / 2;
    rs_init(s, t);
    for (int j = 0; j < size; j++) {
        int b = (int)((i >> (size - 1 - j)) & 1ull);
        if (b == 0 && !((t & 1) && j == size - 1)) {
            if (x == 2 * j) s.J |= bit(2 * j + 1);
            if (x == 2 * j + 1) s.J |= bit(2 * j);
        }
    }
    for (int j = 0; j < size; j++) {
        int b = (int)((i >> (size - 1 - j)) & 1ull);
        if ((t & 1) && j == size - 1) {
            if (b) rs_shrink(s, bit(t - 1), 0, false);
            continue;
        }
        if (b) rs_shrink(s, bit(2 * j) | bit(2 * j + 1), 0, false);
    }
}

// projection by P (innerprod.c:98-116): number of 2^-1/2 factors, -1 if 0
DEV int rs_project(RState &s, const DevProjector &P) {
    int sexp = 0;
    for (int g = 0; g < P.nstabs; g++) {
        int r = rs_measure_pauli(s, P.phase[g] & 3, P.zs[g], P.xs[g]);
        if (r == 0) return -1;
        if (r == 2) sexp++;
    }
    return sexp;
}


// Reduced row echelon basis of the active rows (basis changes of the form,
// so the state is unchanged), then rows and columns of J and the D planes
// re-indexed so that row a sits at its pivot column.
DEV void to_record(const RState &r, int sexp, uint64_t *rec) {
    const int x = lane_id();
    AState s;
    to_astate(r, s);
    int piv = 0;
    u64 Pm = 0;
    for (int a = 0; a < r.k; a++) {
        const u64 Ga = rl64(s.G, a);
        const int pc = __ffsll((long long)Ga) - 1;
        if (x == a) piv = pc;
        Pm |= bit(pc);
        const u64 Sp = ballot(x < r.k && x != a && ((s.G >> pc) & 1ull));
        int dummy;
        as_basis_change<false>(s, a, Sp, true, &dummy);
    }
    // D planes: bit a -> bit piv_a
    const bool act = x < r.k;
    const u64 D1 = wave_or((act && (s.d & 2)) ? bit(piv) : 0ull);
    const u64 D2 = wave_or((act && (s.d & 4)) ? bit(piv) : 0ull);
    // J columns: bit b -> bit piv_b (for active b); rows scattered below
    u64 Jp = 0;
    for (int b = 0; b < r.k; b++) {
        const int pb = rl32(piv, b);
        Jp |= ((s.J >> b) & 1ull) << pb;
    }
    if (x == 0) {
        rec[0] = (uint64_t)r.k | ((uint64_t)r.n << 8) | ((uint64_t)(s.Q & 7) << 16) | ((uint64_t)sexp << 24);
        rec[1] = s.h;
        rec[2] = Pm;
        rec[3] = D1;
        rec[4] = D2;
    }
    // rows of non-pivot columns are never active: clear them, then scatter
    if (!((Pm >> x) & 1ull)) {
        rec[5 + x] = 0;
        rec[69 + x] = 0;
    }
    if (act) {
        rec[5 + piv] = s.G;
        rec[69 + piv] = Jp;
    }
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256)
k_prep_states(int exact, int t, const uint64_t *__restrict__ L, int k, const uint64_t *__restrict__ idx,
              int count, uint64_t *__restrict__ out) {
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (wave >= count) return;
    RState s;
    if (exact) rs_prepH(s, idx[wave], t);
    else rs_prepL(s, idx[wave], t, L, k);
    store_rstate(s, out + (size_t)(3 * t + 4) * wave);
}

extern "C" __global__ void __launch_bounds__(256)
k_random_states(int n, const uint64_t *__restrict__ thr, uint64_t seed, long long first, int count,
                DevProjector P, int use_proj, uint64_t *__restrict__ out, int32_t *__restrict__ sexp,
                int *__restrict__ err) {
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (wave >= count) return;
    RState s;
    if (!rs_random_state(s, n, thr, seed, (uint64_t)(first + wave))) {
        if (lane_id() == 0) atomicOr(err, 1);
        return;
    }
    int se = 0;
    if (use_proj) se = rs_project(s, P);
    if (out) store_rstate(s, out + (size_t)(3 * n + 4) * wave);
    if (sexp && lane_id() == 0) sexp[wave] = se;
}

// The record of Pi prepL(x) built directly in reduced row echelon form,
// rows at their pivot lanes: the product state (|+> on supp(x), |0>
// elsewhere) has act = supp(x), G_a = e_a, h = 0 and no form.  Each generator
// i^m Z(zeta) X(xi) (the cases of measurePauli, stabilizer.c:827-959) works
// on the coordinates c = xi & act of xi in that basis and the residual
// rho = xi + sum_{a in c} G_a (zero exactly when xi is in the span):
//   w = 2m + 4 zeta.h + sum_{a in c} D_a + 4 #edges(J inside c) (mod 8),
//   eta_b = (J c)_b + zeta.G_b on act (J_bb = [D_b in {2,6}]);
//  rho = 0, w in {0,4}: the constraint eta . z = w/4 -- nothing (eta empty,
//    w = 0), zero (eta empty, w = 4), else an elimination of min(eta);
//  rho = 0, w in {2,6}: a factor 2^-1/2 w^{+-1}: Q += s, D -= 2s and J ^= eta
//    eta^T on eta (s = 1 for w = 2, -1 for 6);
//  rho != 0: a new variable at p = min(rho) with G_p = rho, D_p =
//    w + 4 zeta.rho, J row and column eta, then the basis change that clears
//    column p from the other rows.
// Returns the number of 2^-1/2 factors, -1 if the projection vanishes.
DEV int rs_project_rref(AState &s, const DevProjector &P) {
    const int x = lane_id();
    int sexp = 0;
    for (int g = 0; g < P.nstabs; g++) {
        const int m = P.phase[g] & 3;
        const u64 zeta = P.zs[g], xi = P.xs[g];
        const u64 c = xi & s.act;
        const bool inA = lane_in(s.act), inC = lane_in(c);
        const u64 rho = xi ^ wave_xor(inC ? s.G : 0ull);
        const u64 vze = ballot(inA && par64(s.G & zeta));
        const u64 d1 = ballot(inC && (s.d & 2)), d2 = ballot(inC && (s.d & 4));
        const u64 edg = ballot(inC && par64(s.J & c & lowmask(x)));
        const int w = (2 * m + 4 * par64(zeta & s.h) + 2 * __popcll(d1) + 4 * __popcll(d2) + 4 * __popcll(edg)) & 7;
        const u64 jb = (s.J & ~bit(x)) | ((u64)((s.d >> 1) & 1) << x);
        const u64 eta = ballot(inA && par64(jb & c)) ^ vze;
        if (rho == 0 && (w & 2) == 0) {
            if (eta == 0) {
                if (w) return -1;
                continue;
            }
            const int i = __builtin_ctzll(eta);
            as_eliminate<false>(s, i, eta & ~bit(i), w >> 2, true);
        } else if (rho == 0) {
            const int sg = w == 2 ? 1 : -1;
            s.Q = (s.Q + sg) & 7;
            if (lane_in(eta)) {
                s.d = (s.d - 2 * sg) & 7;
                s.J ^= eta;
            }
        } else {
            const int p = __builtin_ctzll(rho);
            const int dp = (w + 4 * par64(zeta & rho)) & 7;
            if (inA) s.J = (s.J & ~bit(p)) | (((eta >> x) & 1ull) << p);
            if (x == p) {
                s.G = rho;
                s.d = dp;
                s.J = eta | ((u64)((dp >> 1) & 1) << p);
            }
            s.act |= bit(p);
            const u64 T = ballot(lane_in(s.act) && x != p && ((s.G >> p) & 1ull));
            if (T) {
                int dummy;
                as_basis_change<false>(s, p, T, true, &dummy);
            }
        }
        sexp++;
    }
    return sexp;
}

// The record of a state already in reduced row echelon form at its pivot
// lanes (the layout to_record produces): J rows restricted to act, the rows
// of non-pivot lanes zero.
DEV void store_rref(const AState &s, int n, int sexp, uint64_t *rec) {
    const int x = lane_id();
    const bool a = lane_in(s.act);
    const u64 D1 = ballot(a && (s.d & 2)), D2 = ballot(a && (s.d & 4));
    if (x == 0) {
        rec[0] = (uint64_t)__popcll(s.act) | ((uint64_t)n << 8) | ((uint64_t)(s.Q & 7) << 16) |
                 ((uint64_t)sexp << 24);
        rec[1] = s.h;
        rec[2] = s.act;
        rec[3] = D1;
        rec[4] = D2;
    }
    rec[5 + x] = a ? s.G : 0ull;
    rec[69 + x] = a ? (s.J & s.act) : 0ull;
}

// Records of projected states, one wave per record g (global index over all
// instances; instance found from the record prefix).
// mode 0: theta_s = random stabilizer state s = first + local index of the
//         instance's Philox stream (seed)
// mode 1: phi_i = prepH(i) (exact) or prepL(i) (|L>), i = local index
// then Pi (innerprod.c:95-116 / :218-241) and the elimination-ready record.
// The instance that holds global index key: the largest i < n with
// off[i] <= key, for nondecreasing off with off[0] <= key.  The whole wave
// searches: each round lane x reads off[lo + x*step] and the ballot's highest
// lane narrows [lo, lo + len) to one step, so up to 4096 instances take two
// dependent loads instead of twelve.  Lane 0 reads off[lo] <= key, true by
// induction; it enters the ballot unconditionally, so no table can drive lo
// below 0.  Every lane must be active.
DEV int find_instance(const long long *__restrict__ off, int n, long long key) {
    const int x = lane_id();
    int lo = 0, len = n;
    while (len > 1) {
        const int step = (len + 63) >> 6;
        const bool ok = x * step < len && off[lo + x * step] <= key;
        const int hl = 63 - __clzll((long long)(ballot(ok) | 1ull));
        lo += hl * step;
        len = min(step, len - hl * step);
    }
    return lo;
}

extern "C" __global__ void __launch_bounds__(256)
k_records(int mode, int ninst, const long long *__restrict__ rec_off, const InstDesc *__restrict__ inst,
          const DevProjector *__restrict__ projs, const uint64_t *__restrict__ thr_table, long long nrec,
          uint64_t *__restrict__ recs, int32_t *__restrict__ sexp_out, int *__restrict__ err) {
    const long long g = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (g >= nrec) return;
    const int in = ninst > 1 ? find_instance(rec_off, ninst, g) : 0;
    const InstDesc &D = inst[in];
    const long long local = g - rec_off[in];
    const int n = D.n;
    uint64_t *rec = recs + (size_t)REC_WORDS * (size_t)g;
    RState s;
    if (mode == 0) {
        if (!rs_random_state(s, n, thr_table + 65 * n, D.seed, (uint64_t)(D.first + local))) {
            if (lane_id() == 0) {
                atomicOr(err, 1);
                sexp_out[g] = -1;
            }
            Rec::store_vanished(rec, n);
            return;
        }
    } else if (D.exact) {
        rs_prepH(s, (u64)local, n);
    } else {
        // |L> records directly in reduced row echelon form
        u64 xv = 0;
        for (int j = 0; j < D.k; j++)
            if (((u64)local >> (D.k - 1 - j)) & 1ull) xv ^= D.L[j];
        AState a;
        a.act = xv & lowmask(n);
        a.G = lane_in(a.act) ? bit(lane_id()) : 0ull;
        a.J = 0;
        a.d = 0;
        a.h = 0;
        a.Q = 0;
        const int se = rs_project_rref(a, projs[in]);
        if (lane_id() == 0) sexp_out[g] = se;
        if (se < 0) Rec::store_vanished(rec, n);
        else store_rref(a, n, se, rec);
        return;
    }
    const int se = rs_project(s, projs[in]);
    if (lane_id() == 0) sexp_out[g] = se;
    if (se < 0) {
        Rec::store_vanished(rec, n);
        return;
    }
    to_record(s, se, rec);
}

// ======================================================================
// The sweep.
namespace {

// Packed term of eps 2^{p/2} w^m: 0 for eps = 0, else 1 + 2 ((-p) 8 + m).
DEV int pack_term(int eps, int np, int m) { return eps ? 1 + 2 * (np * 8 + (m & 7)) : 0; }

// <phi_j | theta'> for phi_j = prepL(j) (a product of |0> and |+>), packed:
// the coordinates outside supp(x_j) are fixed to 0.  On pivot columns this
// fixes the variable itself (x_a = h_a + z_a), all at once; the remaining
// zero coordinates (nonact = non-pivot columns of [0, n)) are restrictions.
template <bool PK>
DEV int ip_phiL(const RecView &rv, u64 lowx, u64 xj, u64 nonact) {
    const u64 F = rv.act & ~xj;
    const u64 B = F & rv.h;
    AState s;
    // Q += sum_{a in B} D_a + 4 #{edges of J inside B}; D_x += 4 |J_x & B|
    // (the two 4 x parity terms share one popcount)
    const u64 JB = rv.J & B;
    const u64 E = ballot(par64(JB & lowx));
    s.Q = rv.Q + 2 * __popcll(rv.D1 & B) + 4 * __popcll((rv.D2 ^ E) & B);
    if (PK) {
        s.J = rv.J ^ ((u64)par64(JB) << 63);
    } else {
        s.J = rv.J;
        s.d = rv.d ^ (par64(JB) << 2);
    }
    s.act = rv.act & xj;
    u64 Zn = nonact & ~xj;
    if (Zn) {
        s.G = rv.G;
        s.h = PK ? rv.h : rv.h ^ wave_xor(lane_in(B) ? rv.G : 0ull);
        bool ok = true;
        while (Zn && ok) {
            const int q = __ffsll((long long)Zn) - 1;
            Zn &= Zn - 1;
            ok = PK ? pk_restrict_coord(s, q, 0, B) : as_restrict<PK>(s, bit(q), 0);
        }
        if (!ok) return 0;
    }
    int pp, mm;
    const int eps = as_expsum<PK>(s, pp, mm);
    return pack_term(eps, rv.k + __popcll(xj) - pp, mm);
}

DEV u64 lrow_combo(const uint64_t *L, int k, u64 j) {
    u64 xv = 0;
    for (int r = 0; r < k; r++)
        if ((j >> (k - 1 - r)) & 1ull) xv ^= L[r];
    return xv;
}

// Wave-uniform constants of one |L> record for phiL_batch: the pivot set
// with bits 60 (the loop's end row) and 63 (kept in act by the loop), its
// part fixed to 1 when outside supp(x_j) (act & h), the D planes, the
// non-pivot columns (at most two), their column ballots over the pivot rows
// (col0 for the lowest, col1 for the other) and h bits (hq bit 0, bit 1),
// Q, and 128 k + 1 with k the record's dimension (W's p field and marker).
struct PhiLConst {
    u64 actR, actH, D1, D2, nonact, col0, col1, h;
    unsigned hq, Q, Kw;
};

// Per-lane tables of a batch: lane c holds x_j of j = jb + c, given as xv
// (with bits 60 and 63 set), and the term-independent part of W:
// m = Q + 2 |D1 & B| in bits 29..31 and 128 (k + |x_j| - n0) + 1 below,
// B = act & h & ~x_j, n0 = |act & x_j| (both extra bits count in n0 and in
// |x_j|).
DEV void phiL_tables(const PhiLConst &C, u64 xv, unsigned &xtlo, unsigned &xthi, unsigned &ttab) {
    xv |= bit(CSE_END_ROW) | bit(63);
    const unsigned n0 = (unsigned)__popcll(C.actR & xv);
    const unsigned px = (unsigned)__popcll(xv);
    const unsigned c1 = (unsigned)__popcll(C.D1 & C.actH & ~xv);
    xtlo = (unsigned)xv;
    xthi = (unsigned)(xv >> 32);
    ttab = ((C.Q + 2u * c1) << 29) + 128u * (px - n0) + C.Kw;
}

// The packed term 16 X + 1 + 2 m of a parked W word (0 stays 0): X from
// bits 7..27, m from bits 29..31 less twice the count of D_a = 6 steps
// (bits 0..6 hold 1 + that count).
DEV int phiL_unpack(int w) {
    const unsigned u = (unsigned)w;
    const unsigned m = ((u >> 29) - 2u * ((u & 0x7fu) - 1u)) & 7u;
    return u ? (int)((((u & 0x0fffffffu) >> 7) << 4) + 1u + (m << 1)) : 0;
}

// nb terms <phi_j | theta'>, j = jb .. jb + nb - 1, as raw W words (see
// phiL_unpack), term c in lane c of the result; lane c of the tables is
// phiL_tables' for j = jb + c.  One asm block computes what ip_phiL<true>
// computes, per term:
//  * fixation of the pivot variables outside supp(x_j): B = act & h & ~x_j;
//    Q += sum_B D + 4 #edges(J inside B) (edges from the parity of J_x & B
//    below x, one ballot; the D part is in the table), D_x += 4 |J_x & B|
//    (bit 63 of row x);
//  * the non-pivot coordinates outside supp(x_j), x_q = 0: with
//    S = col_q & act the constraint S . z = h_q + |col_q & B| mod 2 is an
//    empty sum (S = 0, right side 1), nothing (S = 0, right side 0), or an
//    elimination of min(S) as a constrained step whose variable leaves the
//    exponent (W += 128) and adds no factor; later columns then see
//    G_b ^= G_i on the rows b in S and h ^= G_i when c = 1;
//  * the elimination loop from W = table + 4 |(E ^ D2) & B| in m, whose p
//    field ends as the packed exponent; 0 for a vanishing sum.
// ALL = false (records with at most two non-pivot columns): the columns are
// col0 (the lowest) and col1 with their h bits in hq, the second one updated
// in scalar registers (col1 ^= S and h_q1 ^= c when G_iq1 = 1).
// ALL = true (any number): every column is a ballot of bit q of working G
// rows (glo, ghi; reset from the record's rows g0 at a term's first
// restriction), its h bit from a working h (s[48:49], reset from h0).
// Fixed registers besides the loop's: m0 the term index (the lane the
// tables are read from and v_writelane writes; s61 keeps the caller's m0),
// s[46:47] x_j (then G_i), s[44:45] B, s[48:49] and s50 the column and h
// bit of the restriction at hand (ALL: s[48:49] the working h), s51 the
// constraint's right side / a temporary, s[52:53] the edge ballot / the
// column, s54 the table's W, s[62:63] the pending non-pivot columns before
// the loop.
// The elimination loop of exact-mode terms, which have many variables that
// nothing couples to.  At a term's first step without a constraint (a C
// step), every active variable whose row meets no active variable (bits
// below 61 of J_x & act, its own bit included) and whose D is in {0,4}
// (bit 62 clear; the stored diagonal bit says the same, but the batch does
// not rely on it) is uncoupled for the rest of the loop: each one is a C
// step (a factor 2, W -= 128), or with D = 4 the sum vanishes.  They leave
// act together and a second copy of the loop goes on without the check.
// The end row and bit 63 are kept out of the batch.
#define CSE_ELIM_LOOP_ISO \
    CSE_ELIM_LOOP_BODY("", "s_branch L_iso%=\n") \
    "L_iso%=:\n\t" \
    "s_and_b32 s67, s59, 0x1fffffff\n\t" \
    "s_bitset1_b32 s67, 30\n\t" \
    "v_and_b32 %[vt0], s58, %[jlo]\n\t" \
    "v_and_b32 %[vt1], s67, %[jhi]\n\t" \
    "v_or_b32 %[vt0], %[vt0], %[vt1]\n\t" \
    "v_cmp_eq_u32_e64 s[64:65], 0, %[vt0]\n\t" \
    "v_cmp_gt_i32_e64 s[66:67], 0, %[jhi]\n\t" \
    "s_and_b64 s[64:65], s[64:65], s[58:59]\n\t" \
    "s_and_b32 s65, s65, 0x6fffffff\n\t" \
    "s_and_b64 s[66:67], s[66:67], s[64:65]\n\t" \
    "s_cbranch_scc1 L_isov%=\n\t" \
    "s_andn2_b64 s[58:59], s[58:59], s[64:65]\n\t" \
    "s_bcnt1_i32_b64 s66, s[64:65]\n\t" \
    "s_lshl_b32 s66, s66, 7\n\t" \
    "s_sub_u32 s60, s60, s66\n\t" \
    "s_branch L_top2%=\n" \
    "L_isov%=:\n\t" \
    "s_mov_b32 s56, 0\n\t" \
    "s_branch L_end%=\n" \
    CSE_ELIM_LOOP_BODY("2", "s_branch L_top2%=\n")

#define PHIL_TERM(LOOP) \
    "v_mov_b32 %[parked], 0\n\t" \
    "v_mov_b32 v40, 0xc0000000\n\t" \
    "v_mov_b32 v41, 0x80000000\n\t" \
    "s_mov_b32 s61, m0\n\t" \
    "s_add_u32 m0, %[nb], -1\n\t" \
    "s_nop 0\n" \
    "L_ip%=:\n\t" \
    "v_readlane_b32 s46, %[xtlo], m0\n\t" \
    "v_readlane_b32 s47, %[xthi], m0\n\t" \
    "v_readlane_b32 s54, %[ttab], m0\n\t" \
    /* fixation */ \
    "s_andn2_b64 s[44:45], %[actH], s[46:47]\n\t" \
    "s_and_b64 s[58:59], %[actR], s[46:47]\n\t" \
    "v_and_b32 %[vt0], s44, %[jlo0]\n\t" \
    "v_and_b32 %[vt1], s45, %[jhi0]\n\t" \
    "v_and_b32 %[vt3], %[vt1], %[lxhi]\n\t" \
    /* (S0 & S1) ^ S2: the parity word of J_x & B below x */ \
    "v_bitop3_b32 %[vt2], %[vt0], %[lxlo], %[vt3] bitop3:0x6a\n\t" \
    "v_xor_b32 %[vt3], %[vt0], %[vt1]\n\t" \
    "v_bcnt_u32_b32 %[vt2], %[vt2], 0\n\t" \
    "v_bcnt_u32_b32 %[vt3], %[vt3], 0\n\t" \
    "v_and_b32 %[vt2], 1, %[vt2]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt2]\n\t" \
    "v_mov_b32 %[jlo], %[jlo0]\n\t" \
    "v_lshl_add_u32 %[jhi], %[vt3], 31, %[jhi0]\n\t" \
    "s_xor_b64 s[52:53], s[52:53], %[D2]\n\t" \
    "s_and_b64 s[52:53], s[52:53], s[44:45]\n\t" \
    "s_bcnt1_i32_b64 s51, s[52:53]\n\t" \
    "s_lshl_b32 s51, s51, 31\n\t" \
    "s_add_u32 s60, s54, s51\n\t" \
    "s_andn2_b64 s[62:63], %[nonact], s[46:47]\n\t" \
    "s_cbranch_scc1 L_rs%=\n" \
    "L_elim%=:\n" \
    LOOP \
    "L_end%=:\n\t" \
    "s_cmp_eq_u32 s56, 60\n\t" \
    "s_cselect_b32 s60, s60, 0\n" \
    "L_emit%=:\n\t" \
    "v_writelane_b32 %[parked], s60, m0\n\t" \
    "s_sub_u32 m0, m0, 1\n\t" \
    "s_cbranch_scc0 L_ip%=\n\t" \
    "s_branch L_done%=\n" \
    "L_zero%=:\n\t" \
    "s_mov_b32 s60, 0\n\t" \
    "s_branch L_emit%=\n"

#define PHIL_RESTRICT2 \
    /* restrictions: s[62:63] = the non-pivot columns outside supp(x_j) */ \
    "L_rs%=:\n\t" \
    "s_ff1_i32_b64 s51, %[nonact]\n\t" \
    "s_bitcmp1_b64 s[62:63], s51\n\t" \
    "s_cbranch_scc0 L_rs1%=\n\t" \
    "s_mov_b64 s[48:49], %[col0]\n\t" \
    "s_and_b32 s50, %[hq], 1\n\t" \
    "s_bitset0_b64 s[62:63], s51\n\t" \
    "s_branch L_rdo%=\n" \
    "L_rs1%=:\n\t" \
    "s_mov_b64 s[48:49], %[col1]\n\t" \
    "s_lshr_b32 s50, %[hq], 1\n\t" \
    "s_mov_b64 s[62:63], 0\n" \
    "L_rdo%=:\n\t" \
    "s_and_b64 s[64:65], s[48:49], s[58:59]\n\t" \
    "s_and_b64 s[52:53], s[48:49], s[44:45]\n\t" \
    "s_bcnt1_i32_b64 s51, s[52:53]\n\t" \
    "s_xor_b32 s51, s51, s50\n\t" \
    "s_cmp_lg_u64 s[64:65], 0\n\t" \
    "s_cbranch_scc0 L_rempty%=\n\t" \
    "s_ff1_i32_b64 s56, s[64:65]\n\t" \
    "s_bitset0_b64 s[58:59], s56\n\t" \
    "v_readlane_b32 s66, %[jlo], s56\n\t" \
    "v_readlane_b32 s67, %[jhi], s56\n\t" \
    /* the second column after this elimination */ \
    "s_bitcmp1_b64 %[col1], s56\n\t" \
    "s_cselect_b64 s[52:53], s[64:65], 0\n\t" \
    "s_cselect_b32 s54, s51, 0\n\t" \
    "s_xor_b64 s[48:49], %[col1], s[52:53]\n\t" \
    "s_lshr_b32 s50, %[hq], 1\n\t" \
    "s_xor_b32 s50, s50, s54\n\t" \
    /* elimination of i over S with c = the constraint's right side */ \
    "s_and_b32 s56, s67, 0xe0000000\n\t" \
    "s_and_b32 s67, s67, 0x1fffffff\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t" \
    "s_nop 1\n\t" \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t" \
    "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t" \
    "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t" \
    "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t" \
    "s_bitcmp1_b32 s56, 30\n\t" \
    "s_cbranch_scc0 L_rnojii%=\n\t" \
    "v_xor_b32 %[vt3], %[vt3], %[vt0]\n" \
    "L_rnojii%=:\n\t" \
    "s_add_u32 s60, s60, 0x80\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc0 L_rnoc%=\n\t" \
    "s_or_b32 s65, s65, 0x80000000\n\t" \
    "s_add_u32 s60, s60, s56\n" \
    "L_rnoc%=:\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n" \
    "L_rnext%=:\n\t" \
    "s_cmp_lg_u64 s[62:63], 0\n\t" \
    "s_cbranch_scc0 L_elim%=\n\t" \
    "s_mov_b64 s[62:63], 0\n\t" \
    "s_branch L_rdo%=\n" \
    "L_rempty%=:\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_mov_b64 s[48:49], %[col1]\n\t" \
    "s_lshr_b32 s50, %[hq], 1\n\t" \
    "s_branch L_rnext%=\n"

// One restriction's column, right side and pivot set: q = the lowest
// pending column, its ballot over the working G rows, c = h_q + |col & B|
// (bit 0 of s51), S = col & act; an empty S branches to EMPTY.
#define PHIL_RHEAD(EMPTY) \
    "s_ff1_i32_b64 s51, s[62:63]\n\t" \
    "s_bitset0_b64 s[62:63], s51\n\t" \
    "s_lshl_b64 s[52:53], 1, s51\n\t" \
    "s_lshr_b64 s[66:67], s[48:49], s51\n\t" \
    "v_and_b32 %[vt0], s52, %[glo]\n\t" \
    "v_and_or_b32 %[vt0], s53, %[ghi], %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t" \
    "s_and_b64 s[46:47], s[52:53], s[44:45]\n\t" \
    "s_bcnt1_i32_b64 s51, s[46:47]\n\t" \
    "s_xor_b32 s51, s51, s66\n\t" \
    "s_and_b64 s[64:65], s[52:53], s[58:59]\n\t" \
    "s_cbranch_scc0 " EMPTY "%=\n\t"

// Any number of restrictions (exact mode).  When a term's first restriction
// has an empty constraint, the working rows are still the record's, and
// every pending column that misses act is empty too and stays so (an
// elimination of i over S only changes rows in S, and G_i has no bit q):
// one lane-parallel pass over the record's column table (lane q: column
// q's ballot over the pivot rows, h_q in bit 63) retires them all, or ends
// the term at zero when one of them has right side 1.
#define PHIL_RESTRICT_ALL \
    "L_rs%=:\n\t" \
    "v_mov_b32 %[glo], %[g0lo]\n\t" \
    "v_mov_b32 %[ghi], %[g0hi]\n\t" \
    "s_mov_b64 s[48:49], %[h0]\n\t" \
    PHIL_RHEAD("L_rfirst0") \
    "s_branch L_relim%=\n" \
    "L_rdo%=:\n\t" \
    PHIL_RHEAD("L_rempty") \
    "L_relim%=:\n\t" \
    "s_ff1_i32_b64 s56, s[64:65]\n\t" \
    "s_bitset0_b64 s[58:59], s56\n\t" \
    "v_readlane_b32 s66, %[jlo], s56\n\t" \
    "v_readlane_b32 s67, %[jhi], s56\n\t" \
    "v_readlane_b32 s46, %[glo], s56\n\t" \
    "v_readlane_b32 s47, %[ghi], s56\n\t" \
    /* elimination of i over S with c = the constraint's right side */ \
    "s_and_b32 s56, s67, 0xe0000000\n\t" \
    "s_and_b32 s67, s67, 0x1fffffff\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t" \
    "s_nop 1\n\t" \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t" \
    "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t" \
    "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t" \
    "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t" \
    /* the G rows of S take G_i (row i becomes 0: it is retired) */ \
    "v_bitop3_b32 %[glo], %[glo], s46, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[ghi], %[ghi], s47, %[vt0] bitop3:0x78\n\t" \
    "s_bitcmp1_b32 s56, 30\n\t" \
    "s_cbranch_scc0 L_rnojii%=\n\t" \
    "v_xor_b32 %[vt3], %[vt3], %[vt0]\n" \
    "L_rnojii%=:\n\t" \
    "s_add_u32 s60, s60, 0x80\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc0 L_rnoc%=\n\t" \
    "s_or_b32 s65, s65, 0x80000000\n\t" \
    "s_add_u32 s60, s60, s56\n\t" \
    "s_xor_b64 s[48:49], s[48:49], s[46:47]\n" \
    "L_rnoc%=:\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n" \
    "L_rnext%=:\n\t" \
    "s_cmp_lg_u64 s[62:63], 0\n\t" \
    "s_cbranch_scc1 L_rdo%=\n\t" \
    "s_branch L_elim%=\n" \
    "L_rempty%=:\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_branch L_rnext%=\n" \
    /* the first restriction is empty: E = the pending columns that miss \
       act (below bit 60), zero if one of them has an odd count over B    \
       and h_q (bit 63), else they all leave */ \
    "L_rfirst0%=:\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_and_b32 s67, s59, 0x0fffffff\n\t" \
    "s_or_b32 s50, s45, 0x80000000\n\t" \
    "v_and_b32 %[vt0], s58, %[ctlo]\n\t" \
    "v_and_b32 %[vt1], s67, %[cthi]\n\t" \
    "v_or_b32 %[vt0], %[vt0], %[vt1]\n\t" \
    "v_cmp_eq_u32_e64 s[64:65], 0, %[vt0]\n\t" \
    "v_and_b32 %[vt0], s44, %[ctlo]\n\t" \
    "v_and_b32 %[vt1], s50, %[cthi]\n\t" \
    "v_bcnt_u32_b32 %[vt0], %[vt0], 0\n\t" \
    "v_bcnt_u32_b32 %[vt0], %[vt1], %[vt0]\n\t" \
    "v_and_b32 %[vt0], 1, %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[66:67], 0, %[vt0]\n\t" \
    "s_and_b64 s[64:65], s[64:65], s[62:63]\n\t" \
    "s_and_b64 s[66:67], s[66:67], s[64:65]\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_andn2_b64 s[62:63], s[62:63], s[64:65]\n\t" \
    "s_branch L_rnext%=\n"

template <bool ALL>
DEV int phiL_batch(unsigned nb, unsigned jlo0, unsigned jhi0, unsigned lxlo, unsigned lxhi, unsigned xtlo,
                   unsigned xthi, unsigned ttab, const PhiLConst &C, u64 g0, u64 colt) {
    int parked;
    unsigned jlo, jhi, vt0, vt1, vt2, vt3;
    if constexpr (!ALL) {
        asm volatile(
            PHIL_TERM(CSE_ELIM_LOOP)
            PHIL_RESTRICT2
            "L_done%=:\n\t"
            "s_mov_b32 m0, s61\n\t"
            "s_nop 1\n\t"
            : [parked] "=&v"(parked), [jlo] "=&v"(jlo), [jhi] "=&v"(jhi), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
              [vt2] "=&v"(vt2), [vt3] "=&v"(vt3)
            : [nb] "s"(nb), [jlo0] "v"(jlo0), [jhi0] "v"(jhi0), [lxlo] "v"(lxlo), [lxhi] "v"(lxhi),
              [xtlo] "v"(xtlo), [xthi] "v"(xthi), [ttab] "v"(ttab), [actR] "s"(C.actR), [actH] "s"(C.actH),
              [D2] "s"(C.D2), [nonact] "s"(C.nonact), [col0] "s"(C.col0), [col1] "s"(C.col1), [hq] "s"(C.hq)
            : "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s56", "s58", "s59",
              "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    } else {
        unsigned glo, ghi;
        asm volatile(
            PHIL_TERM(CSE_ELIM_LOOP_ISO)
            PHIL_RESTRICT_ALL
            "L_done%=:\n\t"
            "s_mov_b32 m0, s61\n\t"
            "s_nop 1\n\t"
            : [parked] "=&v"(parked), [jlo] "=&v"(jlo), [jhi] "=&v"(jhi), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
              [vt2] "=&v"(vt2), [vt3] "=&v"(vt3), [glo] "=&v"(glo), [ghi] "=&v"(ghi)
            : [nb] "s"(nb), [jlo0] "v"(jlo0), [jhi0] "v"(jhi0), [lxlo] "v"(lxlo), [lxhi] "v"(lxhi),
              [xtlo] "v"(xtlo), [xthi] "v"(xthi), [ttab] "v"(ttab), [actR] "s"(C.actR), [actH] "s"(C.actH),
              [D2] "s"(C.D2), [nonact] "s"(C.nonact), [g0lo] "v"((unsigned)g0), [g0hi] "v"((unsigned)(g0 >> 32)),
              [h0] "s"(C.h), [ctlo] "v"((unsigned)colt), [cthi] "v"((unsigned)(colt >> 32))
            : "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s56", "s58", "s59",
              "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    }
    return parked;
}

// nb |H^t> terms <prepH(j) | theta'>, j = jb .. jb + nb - 1, as raw W words
// (see phiL_unpack), term c in lane c, the exact-mode kernel's loop.  Pair q
// (qubits 2q, 2q+1) is controlled by bit size-1-q of j (stateprep.c:36-81):
// bit 1 makes it |00> + |11>, the constraint x_2q + x_2q+1 = 0; bit 0 the
// four-term state, the coupling 4 x_2q x_2q+1.  For odd t the last "pair" is
// (t-1, t) with no qubit t: its constraint is x_t-1 = 0 (column t and h_t are
// zero), and it has no coupling (pmask leaves its bit out).  Per term:
//  * working copies of the J rows (D in bits 61..63, the end row), the G
//    rows, h and act; W = wbase - 128 #constraints, wbase = Q in m and
//    128 (k + t - n0) + 1 below;
//  * the constraints, set bits of j from the lowest: restrictions as in
//    PHIL_RESTRICT_ALL, with the column a ballot of the parity of the G rows
//    on the pair and c = the parity of h on it;
//  * the couplings, clear pair bits from the lowest: cu, cv the ballots of
//    bits u, v of the G rows (on act), Q += 4 h_u h_v, D += 4 on
//    (cv & hu) ^ (cu & hv) ^ (cu & cv), J ^= cv on cu and J ^= cu on cv;
//  * the loop: with the batch of uncoupled variables in the exact-mode
//    kernel (ALL), plain in the sampled one.
// Registers as phiL_batch, with s46 j's pair bits (then cv), s50 u = 2q, s51
// the pair's bit, then c, s[52:53] the pair's mask, then the column or cu,
// s62 the pending pair bits.
// a wave-uniform value in a scalar register
DEV unsigned uni32(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

#define PHIH_BODY(LOOP) \
    "v_mov_b32 v40, 0xc0000000\n\t" \
    "v_mov_b32 v41, 0x80000000\n\t" \
    "s_mov_b32 s61, m0\n\t" \
    "s_mov_b32 m0, %[m0s]\n" \
    "L_ip%=:\n\t" \
    "s_add_u32 s46, %[jb], m0\n\t" \
    "s_and_b32 s46, s46, %[smask]\n\t" \
    "s_bcnt1_i32_b32 s47, s46\n\t" \
    "s_lshl_b32 s47, s47, 7\n\t" \
    "s_sub_u32 s60, %[wbase], s47\n\t" \
    "v_mov_b32 %[jlo], %[jlo0]\n\t" \
    "v_mov_b32 %[jhi], %[jhi0]\n\t" \
    "v_mov_b32 %[glo], %[g0lo]\n\t" \
    "v_mov_b32 %[ghi], %[g0hi]\n\t" \
    "s_mov_b64 s[48:49], %[h0]\n\t" \
    "s_mov_b64 s[58:59], %[actR]\n\t" \
    "s_mov_b32 s62, s46\n" \
    /* constraints */ \
    "L_hr%=:\n\t" \
    "s_cmp_eq_u32 s62, 0\n\t" \
    "s_cbranch_scc1 L_hc%=\n\t" \
    "s_ff1_i32_b32 s51, s62\n\t" \
    "s_bitset0_b32 s62, s51\n\t" \
    "s_sub_u32 s50, %[sm1], s51\n\t" \
    "s_lshl_b32 s50, s50, 1\n\t" \
    "s_lshl_b64 s[52:53], 3, s50\n\t" \
    "s_and_b64 s[66:67], s[48:49], s[52:53]\n\t" \
    "s_bcnt1_i32_b64 s51, s[66:67]\n\t" \
    "v_and_b32 %[vt0], s52, %[glo]\n\t" \
    "v_and_or_b32 %[vt0], s53, %[ghi], %[vt0]\n\t" \
    "v_bcnt_u32_b32 %[vt0], %[vt0], 0\n\t" \
    "v_and_b32 %[vt0], 1, %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t" \
    "s_and_b64 s[64:65], s[52:53], s[58:59]\n\t" \
    "s_cbranch_scc0 L_hempty%=\n\t" \
    "s_ff1_i32_b64 s56, s[64:65]\n\t" \
    "s_bitset0_b64 s[58:59], s56\n\t" \
    "v_readlane_b32 s66, %[jlo], s56\n\t" \
    "v_readlane_b32 s67, %[jhi], s56\n\t" \
    "v_readlane_b32 s46, %[glo], s56\n\t" \
    "v_readlane_b32 s47, %[ghi], s56\n\t" \
    "s_and_b32 s56, s67, 0xe0000000\n\t" \
    "s_and_b32 s67, s67, 0x1fffffff\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t" \
    "s_nop 1\n\t" \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t" \
    "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t" \
    "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t" \
    "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t" \
    "v_bitop3_b32 %[glo], %[glo], s46, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[ghi], %[ghi], s47, %[vt0] bitop3:0x78\n\t" \
    "s_bitcmp1_b32 s56, 30\n\t" \
    "s_cbranch_scc0 L_hnojii%=\n\t" \
    "v_xor_b32 %[vt3], %[vt3], %[vt0]\n" \
    "L_hnojii%=:\n\t" \
    "s_add_u32 s60, s60, 0x80\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc0 L_hnoc%=\n\t" \
    "s_or_b32 s65, s65, 0x80000000\n\t" \
    "s_add_u32 s60, s60, s56\n\t" \
    "s_xor_b64 s[48:49], s[48:49], s[46:47]\n" \
    "L_hnoc%=:\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n\t" \
    "s_branch L_hr%=\n" \
    "L_hempty%=:\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_branch L_hr%=\n" \
    /* couplings */ \
    "L_hc%=:\n\t" \
    "s_add_u32 s46, %[jb], m0\n\t" \
    "s_andn2_b32 s62, %[pmask], s46\n" \
    "L_hcl%=:\n\t" \
    "s_cmp_eq_u32 s62, 0\n\t" \
    "s_cbranch_scc1 L_elim%=\n\t" \
    "s_ff1_i32_b32 s51, s62\n\t" \
    "s_bitset0_b32 s62, s51\n\t" \
    "s_sub_u32 s50, %[sm1], s51\n\t" \
    "s_lshl_b32 s50, s50, 1\n\t" \
    "s_lshl_b64 s[52:53], 1, s50\n\t" \
    "s_lshl_b64 s[46:47], 2, s50\n\t" \
    "v_and_b32 %[vt0], s52, %[glo]\n\t" \
    "v_and_or_b32 %[vt0], s53, %[ghi], %[vt0]\n\t" \
    "v_and_b32 %[vt1], s46, %[glo]\n\t" \
    "v_and_or_b32 %[vt1], s47, %[ghi], %[vt1]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[46:47], 0, %[vt1]\n\t" \
    "s_lshr_b64 s[66:67], s[48:49], s50\n\t" \
    "s_and_b64 s[52:53], s[52:53], s[58:59]\n\t" \
    "s_and_b64 s[46:47], s[46:47], s[58:59]\n\t" \
    /* Q += 4 h_u h_v */ \
    "s_and_b32 s67, s66, 3\n\t" \
    "s_cmp_eq_u32 s67, 3\n\t" \
    "s_cselect_b32 s67, 0x80000000, 0\n\t" \
    "s_add_u32 s60, s60, s67\n\t" \
    /* lin = (cv & -h_u) ^ (cu & -h_v) ^ (cu & cv) */ \
    "s_bitcmp1_b32 s66, 0\n\t" \
    "s_cselect_b64 s[64:65], s[46:47], 0\n\t" \
    "s_bitcmp1_b32 s66, 1\n\t" \
    "s_cselect_b64 s[66:67], s[52:53], 0\n\t" \
    "s_xor_b64 s[64:65], s[64:65], s[66:67]\n\t" \
    "s_and_b64 s[66:67], s[52:53], s[46:47]\n\t" \
    "s_xor_b64 s[64:65], s[64:65], s[66:67]\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[52:53]\n\t" \
    "v_cndmask_b32_e64 %[vt1], 0, -1, s[46:47]\n\t" \
    "v_cndmask_b32_e64 %[vt2], 0, -1, s[64:65]\n\t" \
    "v_and_b32 %[vt2], 0x80000000, %[vt2]\n\t" \
    "v_xor_b32 %[jhi], %[jhi], %[vt2]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s46, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s47, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s52, %[vt1] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s53, %[vt1] bitop3:0x78\n\t" \
    "s_branch L_hcl%=\n" \
    "L_elim%=:\n" \
    LOOP \
    "L_end%=:\n\t" \
    "s_cmp_eq_u32 s56, 60\n\t" \
    "s_cselect_b32 s60, s60, 0\n" \
    "L_emit%=:\n\t" \
    "v_writelane_b32 %[parked], s60, m0\n\t" \
    "s_add_u32 m0, m0, 1\n\t" \
    "s_cmp_lt_u32 m0, %[nb]\n\t" \
    "s_cbranch_scc1 L_ip%=\n\t" \
    "s_branch L_done%=\n" \
    "L_zero%=:\n\t" \
    "s_mov_b32 s60, 0\n\t" \
    "s_branch L_emit%=\n" \
    "L_done%=:\n\t" \
    "s_mov_b32 m0, s61\n\t" \
    "s_nop 1\n\t"

// One term's own pair operation for bit B of j in a group of 8 (PHIH_GROUP):
// a constraint when the bit is set, a coupling when it is clear and a pair
// bit (of pmask), nothing otherwise; SH is 2 q for the bit's pair q.  The
// same register use and arithmetic as PHIH_BODY's loops, with the pair's
// masks shifted by a shift computed once per call instead of per operation.
#define PHIH_OWN_BIT(B, SH) \
    "s_bitcmp1_b32 s62, " #B "\n\t" \
    "s_cbranch_scc0 L_gc" #B "%=\n\t" \
    "s_lshl_b64 s[52:53], 3, " SH "\n\t" \
    "s_and_b64 s[66:67], s[48:49], s[52:53]\n\t" \
    "s_bcnt1_i32_b64 s51, s[66:67]\n\t" \
    "v_and_b32 %[vt0], s52, %[glo]\n\t" \
    "v_and_or_b32 %[vt0], s53, %[ghi], %[vt0]\n\t" \
    "v_bcnt_u32_b32 %[vt0], %[vt0], 0\n\t" \
    "v_and_b32 %[vt0], 1, %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t" \
    "s_and_b64 s[64:65], s[52:53], s[58:59]\n\t" \
    "s_cbranch_scc0 L_ge" #B "%=\n\t" \
    "s_ff1_i32_b64 s56, s[64:65]\n\t" \
    "s_bitset0_b64 s[58:59], s56\n\t" \
    "v_readlane_b32 s66, %[jlo], s56\n\t" \
    "v_readlane_b32 s67, %[jhi], s56\n\t" \
    "v_readlane_b32 s46, %[glo], s56\n\t" \
    "v_readlane_b32 s47, %[ghi], s56\n\t" \
    "s_and_b32 s56, s67, 0xe0000000\n\t" \
    "s_and_b32 s67, s67, 0x1fffffff\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t" \
    "s_nop 1\n\t" \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t" \
    "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t" \
    "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t" \
    "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t" \
    "v_bitop3_b32 %[glo], %[glo], s46, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[ghi], %[ghi], s47, %[vt0] bitop3:0x78\n\t" \
    "s_bitcmp1_b32 s56, 30\n\t" \
    "s_cbranch_scc0 L_gnojii" #B "%=\n\t" \
    "v_xor_b32 %[vt3], %[vt3], %[vt0]\n" \
    "L_gnojii" #B "%=:\n\t" \
    "s_add_u32 s60, s60, 0x80\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc0 L_gnoc" #B "%=\n\t" \
    "s_or_b32 s65, s65, 0x80000000\n\t" \
    "s_add_u32 s60, s60, s56\n\t" \
    "s_xor_b64 s[48:49], s[48:49], s[46:47]\n" \
    "L_gnoc" #B "%=:\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n\t" \
    "s_branch L_gn" #B "%=\n" \
    "L_ge" #B "%=:\n\t" \
    "s_bitcmp1_b32 s51, 0\n\t" \
    "s_cbranch_scc1 L_zero%=\n\t" \
    "s_branch L_gn" #B "%=\n" \
    "L_gc" #B "%=:\n\t" \
    "s_bitcmp1_b32 %[pmask], " #B "\n\t" \
    "s_cbranch_scc0 L_gn" #B "%=\n\t" \
    "s_lshl_b64 s[52:53], 1, " SH "\n\t" \
    "s_lshl_b64 s[46:47], 2, " SH "\n\t" \
    "v_and_b32 %[vt0], s52, %[glo]\n\t" \
    "v_and_or_b32 %[vt0], s53, %[ghi], %[vt0]\n\t" \
    "v_and_b32 %[vt1], s46, %[glo]\n\t" \
    "v_and_or_b32 %[vt1], s47, %[ghi], %[vt1]\n\t" \
    "v_cmp_ne_u32_e64 s[52:53], 0, %[vt0]\n\t" \
    "v_cmp_ne_u32_e64 s[46:47], 0, %[vt1]\n\t" \
    "s_lshr_b64 s[66:67], s[48:49], " SH "\n\t" \
    "s_and_b64 s[52:53], s[52:53], s[58:59]\n\t" \
    "s_and_b64 s[46:47], s[46:47], s[58:59]\n\t" \
    "s_and_b32 s67, s66, 3\n\t" \
    "s_cmp_eq_u32 s67, 3\n\t" \
    "s_cselect_b32 s67, 0x80000000, 0\n\t" \
    "s_add_u32 s60, s60, s67\n\t" \
    "s_bitcmp1_b32 s66, 0\n\t" \
    "s_cselect_b64 s[64:65], s[46:47], 0\n\t" \
    "s_bitcmp1_b32 s66, 1\n\t" \
    "s_cselect_b64 s[66:67], s[52:53], 0\n\t" \
    "s_xor_b64 s[64:65], s[64:65], s[66:67]\n\t" \
    "s_and_b64 s[66:67], s[52:53], s[46:47]\n\t" \
    "s_xor_b64 s[64:65], s[64:65], s[66:67]\n\t" \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[52:53]\n\t" \
    "v_cndmask_b32_e64 %[vt1], 0, -1, s[46:47]\n\t" \
    "v_cndmask_b32_e64 %[vt2], 0, -1, s[64:65]\n\t" \
    "v_and_b32 %[vt2], 0x80000000, %[vt2]\n\t" \
    "v_xor_b32 %[jhi], %[jhi], %[vt2]\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s46, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s47, %[vt0] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jlo], %[jlo], s52, %[vt1] bitop3:0x78\n\t" \
    "v_bitop3_b32 %[jhi], %[jhi], s53, %[vt1] bitop3:0x78\n" \
    "L_gn" #B "%=:\n\t"

// PHIH_BODY for the terms of one group of 8 in an aligned batch: bits 3 and
// up of j are the group's (applied by phiH_prefix), so a term's own
// operations are those of bits 0..2, unrolled in bit order (the order of
// pair operations does not change a term's value).  s62 holds the term's
// set bits among them.
#define PHIH_GROUP_BODY(LOOP) \
    "v_mov_b32 v40, 0xc0000000\n\t" \
    "v_mov_b32 v41, 0x80000000\n\t" \
    "s_mov_b32 s61, m0\n\t" \
    "s_mov_b32 m0, %[m0s]\n" \
    "L_ip%=:\n\t" \
    "s_add_u32 s46, %[jb], m0\n\t" \
    "s_and_b32 s62, s46, %[smask]\n\t" \
    "s_bcnt1_i32_b32 s47, s62\n\t" \
    "s_lshl_b32 s47, s47, 7\n\t" \
    "s_sub_u32 s60, %[wbase], s47\n\t" \
    "v_mov_b32 %[jlo], %[jlo0]\n\t" \
    "v_mov_b32 %[jhi], %[jhi0]\n\t" \
    "v_mov_b32 %[glo], %[g0lo]\n\t" \
    "v_mov_b32 %[ghi], %[g0hi]\n\t" \
    "s_mov_b64 s[48:49], %[h0]\n\t" \
    "s_mov_b64 s[58:59], %[actR]\n\t" \
    PHIH_OWN_BIT(0, "%[sh0]") \
    PHIH_OWN_BIT(1, "%[sh1]") \
    PHIH_OWN_BIT(2, "%[sh2]") \
    "L_elim%=:\n" \
    LOOP \
    "L_end%=:\n\t" \
    "s_cmp_eq_u32 s56, 60\n\t" \
    "s_cselect_b32 s60, s60, 0\n" \
    "L_emit%=:\n\t" \
    "v_writelane_b32 %[parked], s60, m0\n\t" \
    "s_add_u32 m0, m0, 1\n\t" \
    "s_cmp_lt_u32 m0, %[nb]\n\t" \
    "s_cbranch_scc1 L_ip%=\n\t" \
    "s_branch L_done%=\n" \
    "L_zero%=:\n\t" \
    "s_mov_b32 s60, 0\n\t" \
    "s_branch L_emit%=\n" \
    "L_done%=:\n\t" \
    "s_mov_b32 m0, s61\n\t" \
    "s_nop 1\n\t"

// Terms jb + m0s .. nb - 1 (a group of 8 of an aligned batch, from the
// group's shared state) into lanes m0s .. nb - 1 of parked, which keeps its
// other lanes; smask and pmask restricted to bits 0..2, sh_b = 2 (sm1 - b).
DEV int phiH_group(unsigned nb, unsigned jb, unsigned jlo0, unsigned jhi0, u64 g0, u64 actR, u64 h0,
                   unsigned wbase, unsigned smask, unsigned pmask, unsigned sh0, unsigned sh1, unsigned sh2,
                   unsigned m0s, int parked) {
    unsigned jlo, jhi, vt0, vt1, vt2, vt3, glo, ghi;
    asm volatile(
        PHIH_GROUP_BODY(CSE_ELIM_LOOP_ISO)
        : [parked] "+&v"(parked), [jlo] "=&v"(jlo), [jhi] "=&v"(jhi), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
          [vt2] "=&v"(vt2), [vt3] "=&v"(vt3), [glo] "=&v"(glo), [ghi] "=&v"(ghi)
        : [nb] "s"(nb), [jb] "s"(jb), [jlo0] "v"(jlo0), [jhi0] "v"(jhi0), [g0lo] "v"((unsigned)g0),
          [g0hi] "v"((unsigned)(g0 >> 32)), [actR] "s"(actR), [h0] "s"(h0), [wbase] "s"(wbase), [smask] "s"(smask),
          [pmask] "s"(pmask), [sh0] "s"(sh0), [sh1] "s"(sh1), [sh2] "s"(sh2), [m0s] "s"(m0s)
        : "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s56", "s58", "s59",
          "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    return parked;
}

// Terms jb + m0s .. jb + nb - 1 into lanes m0s .. nb - 1 of parked, which
// keeps its other lanes (a batch can run as several calls, each from its own
// shared state).
template <bool ALL>
DEV int phiH_batch(unsigned nb, unsigned jb, unsigned jlo0, unsigned jhi0, u64 g0, u64 actR, u64 h0,
                   unsigned wbase, unsigned smask, unsigned sm1, unsigned pmask, unsigned m0s, int parked) {
    unsigned jlo, jhi, vt0, vt1, vt2, vt3, glo, ghi;
    if constexpr (ALL) {
        asm volatile(
            PHIH_BODY(CSE_ELIM_LOOP_ISO)
            : [parked] "+&v"(parked), [jlo] "=&v"(jlo), [jhi] "=&v"(jhi), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
              [vt2] "=&v"(vt2), [vt3] "=&v"(vt3), [glo] "=&v"(glo), [ghi] "=&v"(ghi)
            : [nb] "s"(nb), [jb] "s"(jb), [jlo0] "v"(jlo0), [jhi0] "v"(jhi0), [g0lo] "v"((unsigned)g0),
              [g0hi] "v"((unsigned)(g0 >> 32)), [actR] "s"(actR), [h0] "s"(h0), [wbase] "s"(wbase), [smask] "s"(smask),
              [sm1] "s"(sm1), [pmask] "s"(pmask), [m0s] "s"(m0s)
            : "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s56", "s58", "s59",
              "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    } else {
        asm volatile(
            PHIH_BODY(CSE_ELIM_LOOP)
            : [parked] "+&v"(parked), [jlo] "=&v"(jlo), [jhi] "=&v"(jhi), [vt0] "=&v"(vt0), [vt1] "=&v"(vt1),
              [vt2] "=&v"(vt2), [vt3] "=&v"(vt3), [glo] "=&v"(glo), [ghi] "=&v"(ghi)
            : [nb] "s"(nb), [jb] "s"(jb), [jlo0] "v"(jlo0), [jhi0] "v"(jhi0), [g0lo] "v"((unsigned)g0),
              [g0hi] "v"((unsigned)(g0 >> 32)), [actR] "s"(actR), [h0] "s"(h0), [wbase] "s"(wbase), [smask] "s"(smask),
              [sm1] "s"(sm1), [pmask] "s"(pmask), [m0s] "s"(m0s)
            : "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s56", "s58", "s59",
              "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    }
    return parked;
}

// The pair operations that every term of a phiH_batch batch shares: those
// of the bits of j outside the batch's varying low bits, applied once per
// batch to the record's rows, as phiH_batch applies them per term (the
// same operations in the same order for these bits: constraints for the
// set bits of cbits, then couplings for the bits of pbits, each from the
// lowest; the order of pair operations does not change a term's value).
// J: packed rows (D in bits 61..63), G: rows, h, act, W as phiH_batch keeps
// them; W comes in as the batch's wbase and leaves less 128 per shared
// constraint plus the operations' terms.  Returns false when a shared
// constraint is empty with right side 1: every term of the batch is zero.
DEV bool phiH_prefix(unsigned cbits, unsigned pbits, unsigned sm1, u64 &J, u64 &G, u64 &h, u64 &act,
                     unsigned &W) {
    const int x = lane_id();
    W -= 128u * (unsigned)__popc(cbits);
    while (cbits) {
        const int b = __builtin_ctz(cbits);
        cbits &= cbits - 1;
        const u64 pm = 3ull << (2 * ((int)sm1 - b));
        const unsigned c = (unsigned)__popcll(h & pm) & 1u;
        const u64 S = ballot(__popcll(G & pm) & 1) & act;
        if (!S) {
            if (c) return false;
            continue;
        }
        const int i = __builtin_ctzll(S);
        act &= ~bit(i);
        const u64 Ji = rl64(J, i), Gi = rl64(G, i);
        const unsigned s56 = (unsigned)(Ji >> 32) & 0xe0000000u;
        const u64 Jc = Ji & ~(7ull << 61);
        const bool inS = (S >> x) & 1ull, inJc = (Jc >> x) & 1ull;
        if (inS) {
            // J ^= Jc, D += D_i + 4 J_xi, G ^= G_i (row i retires)
            const u64 Jn = J ^ Jc;
            const unsigned hi = (unsigned)(Jn >> 32) + (s56 ^ (inJc ? 0x80000000u : 0u));
            J = (Jn & 0xffffffffull) | ((u64)hi << 32);
            G ^= Gi;
        }
        // M = Jc, or Jc ^ S when D_i is in {2,6}
        const bool inM = inJc != (((s56 >> 30) & 1u) && inS);
        W += 0x80u;
        u64 Sx = S;
        if (c) {
            Sx |= bit(63);
            W += s56;
            h ^= Gi;
        }
        if (inM) J ^= Sx;
    }
    while (pbits) {
        const int b = __builtin_ctz(pbits);
        pbits &= pbits - 1;
        const int u = 2 * ((int)sm1 - b);
        const u64 cu = ballot((G >> u) & 1ull) & act, cv = ballot((G >> (u + 1)) & 1ull) & act;
        const unsigned hh = (unsigned)(h >> u) & 3u;
        if (hh == 3u) W += 0x80000000u;
        const u64 lin = ((hh & 1u) ? cv : 0ull) ^ ((hh & 2u) ? cu : 0ull) ^ (cu & cv);
        if ((lin >> x) & 1ull) J ^= 1ull << 63;
        if ((cu >> x) & 1ull) J ^= cv;
        if ((cv >> x) & 1ull) J ^= cu;
    }
    return true;
}

// <phi_j | theta'> for phi_j = prepH(j), packed: pair constraints
// x_{2q} = x_{2q+1} (bit 1), couplings 4 x_{2q} x_{2q+1} (bit 0), last qubit
// of odd t fixed to 0 on bit 1 (stateprep.c:36-81, as a coordinate-form
// state).
template <bool PK>
DEV int ip_phiH(const RecView &rv, u64 j, int t) {
    AState s;
    s.G = rv.G;
    s.J = rv.J;
    if (!PK) s.d = rv.d;
    s.act = rv.act;
    s.h = rv.h;
    s.Q = rv.Q;
    const int size = (t + 1) / 2;
    int k2 = t;
    for (int q = 0; q < size; q++) {
        const int b = (int)((j >> (size - 1 - q)) & 1ull);
        if (!b) continue;
        u64 xi = ((t & 1) && q == size - 1) ? bit(t - 1) : (bit(2 * q) | bit(2 * q + 1));
        k2--;
        if (!as_restrict<PK>(s, xi, 0)) return 0;
    }
    for (int q = 0; q < size; q++) {
        const int b = (int)((j >> (size - 1 - q)) & 1ull);
        if (b || ((t & 1) && q == size - 1)) continue;
        as_add_coupling<PK>(s, 2 * q, 2 * q + 1);
    }
    int pp, mm;
    const int eps = as_expsum<PK>(s, pp, mm);
    return pack_term(eps, rv.k + k2 - pp, mm);
}


// item -> instance by binary search on the item prefix
// Record r and chunk c of local item l of an exact-mode instance: block g is
// the last with tri_before(g) <= l (the root of the quadratic, then an
// integer correction of at most a step or two), then the item within it.
DEV void tri_item(long long l, long long ch, long long nch, long long &r, long long &c) {
    const double a = 2.0 * (double)nch + 1.0;
    long long g = (long long)((a - sqrt(a * a - 8.0 * (double)l / (double)ch)) * 0.5);
    if (g < 0) g = 0;
    if (g > nch - 1) g = nch - 1;
    while (g < nch - 1 && tri_before(ch, nch, g + 1) <= l) g++;
    while (g > 0 && tri_before(ch, nch, g) > l) g--;
    const long long off = l - tri_before(ch, nch, g), per = nch - g;
    r = g * ch + off / per;
    c = g + off % per;
}

// FP32 value of 2^{P/2} e^{i pi m/4}
DEV void term_f32(int P, int m, float &re, float &im) {
    const float r = 0.70710678118654752f;
    const int mm = m & 7;
    const float c = (mm == 0) ? 1.f : (mm == 4) ? -1.f : (mm == 2 || mm == 6) ? 0.f : (mm == 1 || mm == 7) ? r : -r;
    const float s = (mm == 2) ? 1.f : (mm == 6) ? -1.f : (mm == 0 || mm == 4) ? 0.f : (mm == 1 || mm == 3) ? r : -r;
    const float mag = ldexpf((P & 1) ? 1.41421356237309505f : 1.f, P >> 1);
    re = mag * c;
    im = mag * s;
}

// Per-lane accumulators: terms are parked one per lane (lane = term count
// mod 64) and folded 64 at a time by all lanes in parallel.
struct LaneAcc {
    long long c0, c1, c2, c3;   // exact Z[w] partial sums, scale 2^32 (mode 0)
    float sr, cr, si, ci;       // Kahan FP32 partial sums
    DEV void clear() {
        c0 = c1 = c2 = c3 = 0;
        sr = cr = si = ci = 0.f;
    }
    DEV void kahan(float vr, float vi) {
        float y = vr - cr;
        float t = sr + y;
        cr = (t - sr) - y;
        sr = t;
        y = vi - ci;
        t = si + y;
        ci = (t - si) - y;
        si = t;
    }
    // + 2^{p/2} w^m in Z[w] coordinates (scale 2^32, -64 <= p <= 0)
    DEV void zomega(int p, int m) {
        const long long v = 1ll << (32 + (p >> 1));
        const int mm = m & 7;
        int qa, qb;
        long long sa, sb;
        if (!(p & 1)) {
            qa = mm; sa = v; qb = 0; sb = 0;
        } else {
            // sqrt2 w^m = w^{m+1} - w^{m+3}
            qa = (mm + 1) & 7; sa = v;
            qb = (mm + 3) & 7; sb = -v;
        }
        if (qa & 4) sa = -sa;
        if (qb & 4) sb = -sb;
        qa &= 3;
        qb &= 3;
        c0 += (qa == 0 ? sa : 0) + (qb == 0 ? sb : 0);
        c1 += (qa == 1 ? sa : 0) + (qb == 1 ? sb : 0);
        c2 += (qa == 2 ? sa : 0) + (qb == 2 ? sb : 0);
        c3 += (qa == 3 ? sa : 0) + (qb == 3 ? sb : 0);
    }
};

DEV long long wave_sum64(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += (long long)shfl_xor64((u64)v, off);
    return (long long)uniform64((u64)v);
}

DEV double wave_sumf64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fold the parked terms (lanes < cnt).  Packed term: 0 for eps = 0, else
// 1 + ((-p) * 8 + m) * 2; in mode 1 the term in lane diag is the diagonal
// one and every other term counts as 2 Re, into histogram bin hoff +
// (-p) * 8 + m.  Returns 1 in a lane whose term is out of range (not folded).
DEV int fold_terms(int packed, int cnt, int mode, int diag, int sexp, LaneAcc &acc, unsigned int *hist,
                   int hoff) {
    const int x = lane_id();
    if (x >= cnt || packed == 0) return 0;
    const int v = packed - 1;
    const int off = mode == 1 && x != diag;
    const int pm = v >> 1;
    // a canonical inner product has -64 <= p <= 0; anything else is an error
    if ((unsigned)pm >= (unsigned)LH_BINS) return 1;
    const int p = -(pm >> 3), m = pm & 7;
    float vr, vi;
    if (mode == 0) {
        // sampled: the per-record Z[w] sums are the result (no histogram)
        acc.zomega(p, m);
        term_f32(p, m, vr, vi);
        acc.kahan(vr, vi);
    } else {
        term_f32(p - sexp, m, vr, vi);
        if (off) {
            acc.kahan(2.f * vr, 0.f);
            atomicAdd(&hist[hoff + pm], 1u);
            atomicAdd(&hist[hoff + ((pm & ~7) | ((8 - m) & 7))], 1u);
        } else {
            acc.kahan(vr, vi);
            atomicAdd(&hist[hoff + pm], 1u);
        }
    }
    return 0;
}

}  // namespace

// Persistent sweep.  mode 0 (sampled): record r is theta_s, all j in [0, chi)
// contribute to per-record Z[w] sums.  mode 1 (exact): record r is theta'_i
// of its instance, j in [i, chi), weight 1 on i == j and 2 Re otherwise, into
// the instance's (P, M) histogram.  MODE is the mode when fixed at compile
// time (0 or 1), -1 for A.mode.  With MODE == 1 a wave keeps its partial
// histogram (over P = p - s, so across records) and its Kahan sums from
// segment to segment of one instance and flushes them to the instance's
// global sums when it moves to another instance and at its exit: one set of
// global atomics per (wave, instance) instead of per segment.
template <bool PK, bool ALL, int MODE>
DEV void sweep_body(const SweepArgs &A) {
    constexpr bool PERSIST = MODE == 1;
    constexpr int HB = MODE == 0 ? 64 : (PERSIST ? HIST_STRIDE : LH_BINS);
    __shared__ unsigned int lhist[4][HB];
    const int mode = MODE >= 0 ? MODE : A.mode;
    const int wv = (int)(threadIdx.x >> 6);
    const int x = lane_id();
    unsigned int *hist = lhist[wv];
    for (int b = x; b < HB; b += 64) hist[b] = 0;
    const u64 lowx = lowmask(x);
    // PERSIST: the instance whose partial sums the wave holds, and those sums
    int cur = -1;
    long long pend = 0;  // terms folded into the partial sums since the last flush
    LaneAcc pacc;
    pacc.clear();
    auto flush_instance = [&]() {
        if (cur < 0) return;
        __threadfence_block();
        const double kr = wave_sumf64((double)pacc.sr - (double)pacc.cr);
        const double ki = wave_sumf64((double)pacc.si - (double)pacc.ci);
        if (x == 0) {
            double *kz = A.kahan + 2 * (size_t)cur;
            atomicAdd(&kz[0], kr);
            atomicAdd(&kz[1], ki);
        }
        for (int b = x; b < HB; b += 64) {
            const unsigned int v = hist[b];
            if (v) {
                atomicAdd((unsigned long long *)&A.hist[(size_t)cur * A.hist_stride + b], (unsigned long long)v);
                hist[b] = 0;
            }
        }
        pacc.clear();
        pend = 0;
    };

    // mode 1: the wave's claimed slice range [s_cur, s_end).  Its first claim
    // is static, slices [g f0, g f0 + f0) of the first static_end for global
    // wave g, so the waves start without queuing on a counter.  The rest,
    // [static_end, nitems), is cut into CSE_NQUEUES equal parts, each with
    // its own counter (128 bytes apart); a wave claims from the part of its
    // block's XCD first (blocks go to the XCDs round robin) and moves on to
    // the next part when that one is drained, so the claims of the whole
    // grid spread over the counters.  Claims are sized by guided
    // self-scheduling from the part's counter read just before the claim, so
    // they shrink as the part drains.  A wave exits after finding every part
    // drained.
    const long long gw = (long long)blockIdx.x * 4 + wv;
    long long s_cur = gw * A.first_take, s_end = s_cur + A.first_take;
    if (s_end > A.static_end) s_end = A.static_end;
    int qx = (int)(blockIdx.x % CSE_NQUEUES), tried = 0;
    long long q0 = A.static_end + (long long)qx * A.part, q1 = q0 + A.part;
    if (q1 > A.nitems) q1 = A.nitems;
    while (true) {
        long long item = 0;
        if (mode == 0) {
            if (x == 0) item = (long long)atomicAdd((unsigned long long *)A.queue, 1ull);
            item = (long long)rl64((u64)item, 0);
            if (item >= A.nitems) break;
        } else {
            while (s_cur >= s_end && tried < CSE_NQUEUES) {
                long long base = 0, take = 0;
                if (x == 0) {
                    // the claim is sized from the part's counter as it is
                    // now (a plain load), so claims shrink as the part
                    // drains; a part seen drained takes no claim
                    unsigned long long *ctr = A.queue + (size_t)qx * CSE_QUEUE_STRIDE;
                    const long long pos =
                        q0 + (long long)__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (pos >= q1) {
                        base = q1;
                    } else {
                        take = (q1 - pos) / (A.waves_per_part * CSE_GSS_DIV);
                        if (take > A.take_max) take = A.take_max;
                        if (take < 1) take = 1;
                        base = q0 + (long long)atomicAdd(ctr, (unsigned long long)take);
                    }
                }
                base = (long long)rl64((u64)base, 0);
                take = (long long)rl64((u64)take, 0);
                if (base < q1) {
                    s_cur = base;
                    s_end = base + take < q1 ? base + take : q1;
                } else {
                    // this part is drained: the next one (a counter only
                    // grows and every claim moves one on, so each part
                    // drains and every wave leaves)
                    tried++;
                    qx = qx + 1 < CSE_NQUEUES ? qx + 1 : 0;
                    q0 = A.static_end + (long long)qx * A.part;
                    q1 = q0 + A.part < A.nitems ? q0 + A.part : A.nitems;
                }
            }
            if (s_cur >= s_end) break;
            item = s_cur;
        }

        const int inst = A.ninst > 1 ? find_instance(A.item_off, A.ninst, item) : 0;
        const InstDesc &D = A.inst[inst];
        const long long local = item - A.item_off[inst];
        long long r, c, nsl = 1;
        if (mode == 0) {
            r = local / D.nchunks;
            c = local % D.nchunks;
        } else {
            tri_item(local, D.chunk, D.nchunks, r, c);
            // one segment: the claimed slices of record r (consecutive in
            // the enumeration, from slice c to the record's last)
            nsl = s_end - s_cur;
            if (nsl > D.nchunks - c) nsl = D.nchunks - c;
            s_cur += nsl;
        }
        long long j0 = c * (long long)D.chunk;
        long long j1 = j0 + nsl * (long long)D.chunk;
        if (j1 > D.chi) j1 = D.chi;
        if (mode == 1 && j0 < r) j0 = r;
        if (j0 >= j1) continue;
        if (PERSIST && inst != cur) {
            flush_instance();
            cur = inst;
        }
        // first per-term output slot of j0
        const long long tbase = mode == 0 ? (D.term_off + r * D.chi + j0)
                                            : (D.term_off + r * D.chi - r * (r - 1) / 2 + (j0 - r));

        const uint64_t *rec = A.recs + (size_t)REC_WORDS * (size_t)(D.rec_off + r);
        RecView rv;
        rv.load(rec);
        if (PK && !rv.vanished) {
            rv.J |= (u64)rv.d << 61;
            // the end row of the elimination loop: J = 0, D = 4
            if (x == CSE_END_ROW) rv.J = 4ull << 61;
        }
        if (rv.vanished) {
            if (A.terms)
                for (long long j = j0 + x; j < j1; j += 64) A.terms[tbase + (j - j0)] = 0;
            continue;
        }

        LaneAcc iacc;
        iacc.clear();
        LaneAcc &acc = PERSIST ? pacc : iacc;
        // instance fields used per term, held in registers
        const int n = D.n;
        int bad = 0;  // per lane: a term outside p in [-64, 0] (never folded)
        // a batch of up to 64 terms, one parked per lane, then folded by all
        // lanes; in mode 1 every term but j = r counts twice (2 Re)
        auto flush = [&](int parked, int nb, long long jb) {
            if (A.terms && x < nb) {
                const int v = parked - 1;
                A.terms[tbase + (jb - j0) + x] = parked ? ((v >> 1) + 1) : 0;
            }
            const int diag = (mode == 1 && r >= jb && r < jb + nb) ? (int)(r - jb) : -1;
            bad |= fold_terms(parked, nb, mode, diag, rv.sexp, acc, hist, PERSIST ? rv.sexp * 8 : 0);
        };
        if (D.exact && PK) {
            // the pair operations the batch shares once, then the whole
            // batch in one asm block
            const unsigned size = (unsigned)((n + 1) / 2);
            const unsigned smask = (1u << size) - 1u;
            const unsigned pmask = smask & ~(unsigned)(n & 1);
            const unsigned wbase = ((unsigned)rv.Q << 29) + 128u * (unsigned)(rv.k + n - __popcll(rv.act)) + 1u;
            const u64 actR = rv.act | bit(CSE_END_ROW) | bit(63);
            for (long long jb = j0; jb < j1; jb += 64) {
                const int nb = (int)(j1 - jb < 64 ? j1 - jb : 64);
                // the bits of j the batch's terms agree on: their pair
                // operations run once, in phiH_prefix
                const unsigned d = (unsigned)(jb ^ (jb + nb - 1));
                const unsigned vary = d ? (2u << (31 - __builtin_clz(d))) - 1u : 0u;
                u64 J = rv.J, G = rv.G, h = rv.h, act = actR;
                unsigned W = wbase;
                int parked = 0;
                if (phiH_prefix((unsigned)jb & smask & ~vary, pmask & ~(unsigned)jb & ~vary, size - 1u, J, G, h,
                                act, W)) {
                    if (ALL && vary == 63u && nb == 64) {
                        // an aligned batch: bits 3..5 are shared by each
                        // group of 8 terms, once per group (not in the
                        // sampled |L> kernel: the group's copy of the
                        // state would cost it VGPRs)
                        for (unsigned g = 0; g < 64u; g += 8u) {
                            u64 J8 = J, G8 = G, h8 = h, act8 = act;
                            unsigned W8 = W;
                            const unsigned jg = (unsigned)jb + g;
                            if (phiH_prefix(jg & smask & 0x38u, pmask & ~jg & 0x38u, size - 1u, J8, G8, h8, act8, W8))
                                parked = phiH_group(uni32(g + 8u), uni32((unsigned)jb), (unsigned)J8,
                                                    (unsigned)(J8 >> 32), G8, uniform64(act8), uniform64(h8),
                                                    uni32(W8), uni32(smask & 7u), uni32(pmask & 7u),
                                                    uni32(2u * (size - 1u)), uni32(2u * (size - 2u)),
                                                    uni32(2u * (size - 3u)), uni32(g), parked);
                        }
                    } else {
                        parked = phiH_batch<ALL>(uni32((unsigned)nb), uni32((unsigned)jb), (unsigned)J,
                                                 (unsigned)(J >> 32), G, uniform64(act), uniform64(h), uni32(W),
                                                 uni32(smask & vary), uni32(size - 1u), uni32(pmask & vary), 0u,
                                                 parked);
                    }
                }
                parked = phiL_unpack(parked);
                flush(parked, nb, jb);
            }
        } else if (D.exact) {
            for (long long jb = j0; jb < j1; jb += 64) {
                const int nb = (int)(j1 - jb < 64 ? j1 - jb : 64);
                int parked = 0;
                for (int c = 0; c < nb; c++) {
                    const int packed = ip_phiH<PK>(rv, (u64)(jb + c), n);
                    if (lane_in(bit(c))) parked = packed;
                }
                flush(parked, nb, jb);
            }
        } else {
            const int k = D.k;
            const u64 nonact = lowmask(n) & ~rv.act;
            if (PK && (ALL || __popcll(nonact) <= 2)) {
                // the whole batch in one asm block
                PhiLConst C;
                C.actR = rv.act | bit(CSE_END_ROW) | bit(63);
                C.actH = rv.act & rv.h;
                C.D1 = rv.D1;
                C.D2 = rv.D2;
                C.nonact = nonact;
                C.h = rv.h;
                // ALL: lane q holds column q's ballot over the pivot rows and
                // h_q in bit 63, for every non-pivot column q
                u64 colt = 0;
                if (ALL) {
                    for (u64 rest = nonact; rest; rest &= rest - 1) {
                        const int q = __builtin_ctzll(rest);
                        const u64 c = ballot((rv.G >> q) & 1ull) & rv.act;
                        if (x == q) colt = c | (((rv.h >> q) & 1ull) << 63);
                    }
                }
                if (!ALL) {
                    const u64 q0m = nonact & (0ull - nonact), q1m = nonact & ~q0m;
                    const int q0 = q0m ? __builtin_ctzll(q0m) : 0, q1 = q1m ? __builtin_ctzll(q1m) : 0;
                    C.col0 = q0m ? (ballot((rv.G >> q0) & 1ull) & rv.act) : 0ull;
                    C.col1 = q1m ? (ballot((rv.G >> q1) & 1ull) & rv.act) : 0ull;
                    C.hq = (unsigned)((q0m ? (rv.h >> q0) & 1ull : 0ull) |
                                      (q1m ? ((rv.h >> q1) & 1ull) << 1 : 0ull));
                }
                C.Q = (unsigned)rv.Q;
                C.Kw = 128u * (unsigned)rv.k + 1u;
                const unsigned jlo0 = (unsigned)rv.J, jhi0 = (unsigned)(rv.J >> 32);
                // x_{jb + c} for batches at multiples of 64: the part of the
                // low six bits of j (lane c, xlow) and that of the others
                // (uniform, xhigh, advanced from batch to batch by the rows
                // whose bits flip: lane b of fliph holds the XOR of rows
                // k-7 .. k-7-b)
                const bool aligned = (j0 & 63) == 0;
                u64 xlow = 0, fliph = 0, xhigh = 0;
                if (aligned) {
                    for (int b = 0; b < 6 && b < k; b++)
                        if ((x >> b) & 1) xlow ^= D.L[k - 1 - b];
                    if (x < k - 6)
                        for (int b = 0; b <= x; b++) fliph ^= D.L[k - 7 - b];
                    xhigh = lrow_combo(D.L, k, (u64)j0);
                }
                for (long long jb = j0; jb < j1; jb += 64) {
                    const int nb = (int)(j1 - jb < 64 ? j1 - jb : 64);
                    u64 xv;
                    if (aligned) {
                        xv = xlow ^ xhigh;
                        // 0 < (jb >> 6) + 1 <= chi / 64 <= 2^24
                        xhigh ^= rl64(fliph, __builtin_ctz((unsigned)((jb >> 6) + 1)));
                    } else {
                        // x_j = sum of the rows r of L with bit k-1-r of j
                        // set, per lane
                        const u64 j = (u64)(jb + x);
                        xv = 0;
                        for (int rr = 0; rr < k; rr++) xv ^= D.L[rr] & (0ull - ((j >> (k - 1 - rr)) & 1ull));
                    }
                    unsigned xtlo, xthi, ttab;
                    phiL_tables(C, xv, xtlo, xthi, ttab);
                    const int parked = phiL_unpack(phiL_batch<ALL>(uni32((unsigned)nb), jlo0, jhi0, (unsigned)lowx,
                                                                   (unsigned)(lowx >> 32), xtlo, xthi, ttab, C,
                                                                   rv.G, colt));
                    flush(parked, nb, jb);
                }
            } else {
                // x_j of |L> incrementally: going from j-1 to j flips bits
                // 0..c of j, c = ctz(j), i.e. rows k-1 .. k-1-c of L; lane c
                // holds their XOR
                u64 xj = lrow_combo(D.L, k, (u64)j0);
                u64 flip = 0;
                if (x < k) {
                    for (int b = 0; b <= x; b++) flip ^= D.L[k - 1 - b];
                }
                for (long long jb = j0; jb < j1; jb += 64) {
                    const int nb = (int)(j1 - jb < 64 ? j1 - jb : 64);
                    int parked = 0;
                    for (int c = 0; c < nb; c++) {
                        const int packed = ip_phiL<PK>(rv, lowx, xj, nonact);
                        if (lane_in(bit(c))) parked = packed;
                        // 0 < j + 1 <= chi <= 2^30
                        xj ^= rl64(flip, __builtin_ctz((unsigned)(jb + c + 1)));
                    }
                    flush(parked, nb, jb);
                }
            }
        }
        if (ballot(bad) && x == 0) atomicOr(A.err, 2);
        if (x == 0) atomicAdd((unsigned long long *)&A.ips[inst], (unsigned long long)(j1 - j0));
        if (PERSIST) {
            pend += j1 - j0;
            if (pend >= A.flush_terms) flush_instance();
            continue;
        }
        // flush the item (the LDS histogram atomics of this wave come first)
        __threadfence_block();
        const double kr = wave_sumf64((double)acc.sr - (double)acc.cr);
        const double ki = wave_sumf64((double)acc.si - (double)acc.ci);
        if (mode == 0) {
            const long long z0 = wave_sum64(acc.c0), z1 = wave_sum64(acc.c1);
            const long long z2 = wave_sum64(acc.c2), z3 = wave_sum64(acc.c3);
            if (x == 0) {
                long long *z = A.zsum + 4 * (size_t)(D.rec_off + r);
                atomicAdd((unsigned long long *)&z[0], (unsigned long long)z0);
                atomicAdd((unsigned long long *)&z[1], (unsigned long long)z1);
                atomicAdd((unsigned long long *)&z[2], (unsigned long long)z2);
                atomicAdd((unsigned long long *)&z[3], (unsigned long long)z3);
                double *kz = A.kahan + 2 * (size_t)(D.rec_off + r);
                atomicAdd(&kz[0], kr);
                atomicAdd(&kz[1], ki);
            }
        } else if (x == 0) {
            double *kz = A.kahan + 2 * (size_t)inst;
            atomicAdd(&kz[0], kr);
            atomicAdd(&kz[1], ki);
        }
        const int shift = rv.sexp * 8;
        for (int b = x; mode == 1 && b < LH_BINS; b += 64) {
            unsigned int v = hist[b];
            if (v) {
                atomicAdd((unsigned long long *)&A.hist[(size_t)inst * A.hist_stride + b + shift],
                          (unsigned long long)v);
                hist[b] = 0;
            }
        }
    }
    if constexpr (PERSIST) {
        // The last flush is merged per block: the waves put their Kahan
        // partials and instance in LDS, and the lowest wave holding each
        // instance adds the block's histograms and partials for it to the
        // global sums.  On a launch of one instance every wave's last flush
        // lands on the same addresses, the two Kahan sums above all; a
        // quarter of the atomics there shortens the end of the launch
        // (C2: 0.67 -> 0.54 ms).
        __shared__ int bcur[4];
        __shared__ double bk[4][2];
        const double kr = wave_sumf64((double)pacc.sr - (double)pacc.cr);
        const double ki = wave_sumf64((double)pacc.si - (double)pacc.ci);
        if (x == 0) {
            bcur[wv] = cur;
            bk[wv][0] = kr;
            bk[wv][1] = ki;
        }
        __syncthreads();
        bool lead = cur >= 0;
        for (int w = 0; w < wv; w++)
            if (bcur[w] == cur) lead = false;
        if (lead) {
            double sr = 0.0, si = 0.0;
            for (int w = wv; w < 4; w++)
                if (bcur[w] == cur) {
                    sr += bk[w][0];
                    si += bk[w][1];
                }
            if (x == 0) {
                double *kz = A.kahan + 2 * (size_t)cur;
                atomicAdd(&kz[0], sr);
                atomicAdd(&kz[1], si);
            }
            for (int b = x; b < HB; b += 64) {
                unsigned long long v = 0;
                for (int w = wv; w < 4; w++)
                    if (bcur[w] == cur) v += lhist[w][b];
                if (v) atomicAdd((unsigned long long *)&A.hist[(size_t)cur * A.hist_stride + b], v);
            }
        }
    }
}

extern "C" __global__ void __launch_bounds__(256) k_sweep(SweepArgs A) { sweep_body<false, false, -1>(A); }

// the same sweep with D packed into the J rows (every instance has n <= 60;
// row 60 is the elimination loop's end row).
// Eight waves per SIMD (at most 64 VGPRs; the SGPRs beyond the budget spill
// to VGPR lanes, nothing to scratch): the elimination loop is bound by
// instruction issue, and the extra wave hides more of its VALU -> SALU
// dependency latency.
extern "C" __global__ void __launch_bounds__(256, 8) k_sweep_packed(SweepArgs A) {
    sweep_body<true, false, 0>(A);
}

// The sampled sweep with the general restriction loop of exact mode: every
// |L> term runs in the asm block whatever the record's number of non-pivot
// columns, every |H^t> term in phiH_batch with the batch of uncoupled
// variables, an aligned batch as eight groups of 8 terms.  Eight waves per
// SIMD (64 VGPRs).  The engine's choice for sampled calls of |H^t> terms
// only, else a tuning alternative to k_sweep_packed (cse_set_sweep_kernel);
// the results are identical.
extern "C" __global__ void __launch_bounds__(256, 8) k_sweep_packed_gen(SweepArgs A) {
    sweep_body<true, true, 0>(A);
}

// The packed sweep of exact mode: exact-mode records (projected |L> states of dimension about |x_i|) have
// about n/2 non-pivot columns, so every |L> term runs in the asm block with
// its general restriction loop, and every |H^t> term in phiH_batch.  The working G
// rows take two more VGPRs: seven waves per SIMD (72 VGPRs), with the
// persistent partial sums and queue state spilled to scratch around the asm
// blocks, once per batch of 64 terms.
extern "C" __global__ void __launch_bounds__(256, CSE_EXACT_BLOCKS_PER_CU) k_sweep_packed_all(SweepArgs A) {
    sweep_body<true, true, 1>(A);
}

// The same exact-mode sweep at eight waves per SIMD (64 VGPRs, more values
// spilled to VGPR lanes and scratch): more latency hidden, more spill
// traffic.  A tuning
// alternative (cse_set_sweep_kernel); large mixed launches gain a few
// percent, launches of many small instances lose a few.
extern "C" __global__ void __launch_bounds__(256, 8) k_sweep_packed_all8(SweepArgs A) {
    sweep_body<true, true, 1>(A);
}

// Histogram of the Hamming weight |x| over the 2^k combinations
// x = sum_j i_j L_j of the rows of L (decompose's Z_L, probability.c:
// 371-410).  Thread g of the grid walks the 2^shift combinations
// i = (g << shift) + (Gray code of s), s = 0 .. 2^shift - 1: bit b of i
// selects row k-1-b, as in the reference's MSB-first binrep.  The grid
// holds exactly 2^(k-shift) threads in full blocks of 256 (k - shift >= 8)
// or one block with the surplus threads idle.
extern "C" __global__ void __launch_bounds__(256)
k_lspace_hist(const uint64_t *__restrict__ L, int k, int shift, unsigned long long *__restrict__ hist) {
    __shared__ unsigned int lh[65];
    __shared__ uint64_t rows[64];
    const int tid = (int)threadIdx.x;
    if (tid < 65) lh[tid] = 0;
    if (tid < k) rows[tid] = L[tid];
    __syncthreads();
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + (uint64_t)tid;
    if ((g >> (k - shift)) == 0) {
        const uint64_t base = g << shift;
        uint64_t x = 0;
        for (int b = shift; b < k; b++)
            if ((base >> b) & 1ull) x ^= rows[k - 1 - b];
        atomicAdd(&lh[__popcll(x)], 1u);
        const uint64_t steps = 1ull << shift;
        for (uint64_t s = 1; s < steps; s++) {
            x ^= rows[k - 1 - __builtin_ctzll(s)];
            atomicAdd(&lh[__popcll(x)], 1u);
        }
    }
    __syncthreads();
    if (tid < 65 && lh[tid]) atomicAdd(&hist[tid], (unsigned long long)lh[tid]);
}

extern "C" __global__ void k_zero64(unsigned long long *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// The per-call sums, counters and flags of a batch (up to 8 buffers of
// 64-bit words) cleared in one launch, grid-stride over each buffer.
struct ZeroList {
    unsigned long long *p[8];
    long long n[8];
    int count;
};

extern "C" __global__ void k_zero_list(ZeroList L) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int b = 0; b < L.count; b++)
        for (long long i = i0; i < L.n[b]; i += stride) L.p[b][i] = 0ull;
}
