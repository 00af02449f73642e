// cse_device.h -- wave64 device routines of the stabilizer-rank inner-product
// engine for gfx950 (MI355X).
//
// Execution model: one wavefront (64 lanes) owns one stabilizer state of
// dimension n <= 64.  Lane x holds row x of G, Gbar and J as uint64 bit rows
// (bit j = column j) and the value D_x in Z8; the scalars (n, k, Q, h, the
// active-row mask) are wave-uniform and live in SGPRs.  GF(2) elimination
// steps are rank-1 updates: the pivot set is a ballot, the pivot row is
// broadcast with v_readlane, and every row updates itself in O(1).  There is
// no MFMA: the work is bitwise.
//
// Two layouts of the same mathematical object are used:
//  * reference layout (RState): active rows are 0..k-1 and row swaps follow
//    the reference's bookkeeping (libcirc/stabilizer/stabilizer.c:500-585,
//    :759-825, :827-959), so a state built on the device equals, row for row,
//    the one the reference builds;
//  * active-mask layout (AState): the active rows are an arbitrary uniform
//    mask, pivots are simply retired (no swaps), and only the quadratic form
//    on active rows is kept.  Every inner product is computed in this layout;
//    the result eps 2^{p/2} w^m is canonical, so it equals the reference's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEV __device__ __forceinline__


typedef unsigned long long u64;

namespace cse {

DEV int lane_id() { return (int)(threadIdx.x & 63); }
DEV u64 bit(int i) { return 1ull << i; }
DEV u64 lowmask(int k) { return k >= 64 ? ~0ull : (bit(k) - 1ull); }
DEV int par64(u64 v) { return __popcll(v) & 1; }

DEV u64 ballot(bool p) { return (u64)__ballot(p); }

DEV int rl32(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
DEV u64 rl64(u64 v, int lane) {
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, lane);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return ((u64)hi << 32) | lo;
}
DEV u64 uniform64(u64 v) { return rl64(v, 0); }

DEV u64 shfl_xor64(u64 v, int m) {
    int lo = __shfl_xor((int)(unsigned)v, m, 64);
    int hi = __shfl_xor((int)(unsigned)(v >> 32), m, 64);
    return ((u64)(unsigned)hi << 32) | (unsigned)lo;
}
// Cross-lane moves within the rows of 16 lanes (DPP, in registers):
// quad_perm [1,0,3,2] (0xB1) and [2,3,0,1] (0x4E), row_ror 4 (0x124) and 8
// (0x128).  Every source lane is inside the row, and the callers run with
// all 64 lanes active.
template <int CTRL> DEV unsigned dpp_mov(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
// XOR of a 32-bit value over each row: the two quad exchanges give every lane
// its quad's XOR, the two rotations fold the row's four quads
DEV unsigned row_xor32(unsigned v) {
    v ^= dpp_mov<0xB1>(v);
    v ^= dpp_mov<0x4E>(v);
    v ^= dpp_mov<0x124>(v);
    v ^= dpp_mov<0x128>(v);
    return v;
}
DEV unsigned row_or32(unsigned v) {
    v |= dpp_mov<0xB1>(v);
    v |= dpp_mov<0x4E>(v);
    v |= dpp_mov<0x124>(v);
    v |= dpp_mov<0x128>(v);
    return v;
}
// XOR of v over all lanes, wave-uniform: rows reduced with DPP, the four
// rows' results (lanes 0, 16, 32, 48) combined in scalar registers
DEV u64 wave_xor(u64 v) {
    const int lo = (int)row_xor32((unsigned)v), hi = (int)row_xor32((unsigned)(v >> 32));
    const unsigned l = (unsigned)(rl32(lo, 0) ^ rl32(lo, 16) ^ rl32(lo, 32) ^ rl32(lo, 48));
    const unsigned h = (unsigned)(rl32(hi, 0) ^ rl32(hi, 16) ^ rl32(hi, 32) ^ rl32(hi, 48));
    return ((u64)h << 32) | l;
}
// OR of v over all lanes, wave-uniform (as wave_xor)
DEV u64 wave_or(u64 v) {
    const int lo = (int)row_or32((unsigned)v), hi = (int)row_or32((unsigned)(v >> 32));
    const unsigned l = (unsigned)(rl32(lo, 0) | rl32(lo, 16) | rl32(lo, 32) | rl32(lo, 48));
    const unsigned h = (unsigned)(rl32(hi, 0) | rl32(hi, 16) | rl32(hi, 32) | rl32(hi, 48));
    return ((u64)h << 32) | l;
}

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct P4 { uint32_t x[4]; };
DEV P4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    P4 o; o.x[0] = c0; o.x[1] = c1; o.x[2] = c2; o.x[3] = c3;
    return o;
}

// Sequential 64-bit draws from the stream (key, stream id, domain): word w
// is output (w & 3) of block w >> 2; a 64-bit draw is words (w, w+1).
struct Stream {
    uint32_t k0, k1, s0, s1, dom;
    uint64_t w;
    P4 blk;
    uint64_t blk_id;
    DEV void init(uint64_t key, uint64_t stream, uint32_t domain) {
        k0 = (uint32_t)key; k1 = (uint32_t)(key >> 32);
        s0 = (uint32_t)stream; s1 = (uint32_t)(stream >> 32);
        dom = domain; w = 0; blk_id = ~0ull;
    }
    DEV uint32_t word() {
        uint64_t b = w >> 2;
        if (b != blk_id) { blk = philox((uint32_t)b, s0, s1, dom, k0, k1); blk_id = b; }
        const int q = (int)(w & 3);
        uint32_t v = q == 0 ? blk.x[0] : q == 1 ? blk.x[1] : q == 2 ? blk.x[2] : blk.x[3];
        w++;
        return v;
    }
    DEV u64 next64() { u64 lo = word(); u64 hi = word(); return lo | (hi << 32); }
};

// ------------------------------------------------------- reference layout
struct RState {
    // per lane
    u64 G, Gb, J;
    int d;
    // uniform
    int n, k, Q;
    u64 h;
};

DEV void rs_init(RState &s, int n) {
    int x = lane_id();
    s.G = x < n ? bit(x) : 0;
    s.Gb = s.G;
    s.J = 0;
    s.d = 0;
    s.n = n; s.k = n; s.Q = 0; s.h = 0;
}

// swap rows i and j of a per-lane row array
DEV u64 swap_rows(u64 v, int i, int j) {
    u64 vi = rl64(v, i), vj = rl64(v, j);
    int x = lane_id();
    return x == i ? vj : (x == j ? vi : v);
}
DEV int swap_rows32(int v, int i, int j) {
    int vi = rl32(v, i), vj = rl32(v, j);
    int x = lane_id();
    return x == i ? vj : (x == j ? vi : v);
}
DEV u64 swap_bits(u64 v, int i, int j) {
    u64 t = ((v >> i) ^ (v >> j)) & 1ull;
    return v ^ ((t << i) | (t << j));
}

// stabilizer.c:500-585.  0 EMPTY, 1 SAME, 2 SUCCESS.
DEV int rs_shrink(RState &s, u64 xi, int alpha, bool lazy) {
    const int x = lane_id();
    u64 S = ballot(x < s.k && par64(s.G & xi));
    int beta = (alpha + __popcll(xi & s.h)) & 1;
    if (S == 0) return 1 - beta;
    int i = 63 - __clzll(S);
    u64 Sp = S & ~bit(i);
    bool inSp = (Sp >> x) & 1;
    u64 Gi = rl64(s.G, i);
    if (inSp) s.G ^= Gi;
    if (!lazy && Sp) {
        // R = I + sum_{a in Sp} e_a e_i^T (sequential updateDJ calls commute)
        u64 Ji = rl64(s.J, i);
        int di = rl32(s.d, i);
        int jxi = (int)((s.J >> i) & 1ull);
        if (inSp) {
            s.d = (s.d + di + 4 * jxi) & 7;
            s.J ^= Ji;
        }
        if ((s.J >> i) & 1ull) s.J ^= Sp;
    }
    if (!lazy && x >= s.k) s.d = 0;  // updateDJ writes D_x = 0 beyond k
    u64 red = wave_xor(inSp ? s.Gb : 0ull);
    if (x == i) s.Gb ^= red;
    const int last = s.k - 1;
    s.G = swap_rows(s.G, i, last);
    s.Gb = swap_rows(s.Gb, i, last);
    if (!lazy) {
        s.d = swap_rows32(s.d, i, last);
        s.J = swap_rows(s.J, i, last);
        s.J = swap_bits(s.J, i, last);
    }
    if (beta) s.h ^= Gi;
    if (!lazy) {
        if (beta) {
            int dl = rl32(s.d, last);
            s.Q = (s.Q + dl) & 7;
            if ((s.J >> last) & 1ull) s.d ^= 4;
        }
    }
    s.k--;
    return 2;
}

// stabilizer.c:759-825
DEV void rs_extend(RState &s, u64 xi) {
    const int x = lane_id();
    u64 S = ballot(x < s.n && par64(xi & s.Gb));
    u64 T = (s.k < s.n) ? (S & ~lowmask(s.k)) : 0ull;
    if (T == 0) return;
    int i = __ffsll((long long)T) - 1;
    u64 Sp = S & ~bit(i);
    bool inSp = (Sp >> x) & 1;
    u64 Gbi = rl64(s.Gb, i);
    u64 red = wave_xor(inSp ? s.G : 0ull);
    if (inSp) s.Gb ^= Gbi;
    if (x == i) s.G ^= red;
    s.G = swap_rows(s.G, i, s.k);
    s.Gb = swap_rows(s.Gb, i, s.k);
    s.k++;
}

// stabilizer.c:827-959 for P = i^m Z(zeta) X(xi).  0 -> norm 0, 1 -> 1, 2 -> 2^-1/2
DEV int rs_measure_pauli(RState &s, int m, u64 zeta, u64 xi) {
    const int x = lane_id();
    const bool act = x < s.k;
    // coordinates of xi in the basis G (G Gbar^T = I over all n rows): xi is
    // in span(G_0 .. G_{k-1}) (xi' = sum of the rows in vxi equals xi)
    // exactly when xi . Gbar_b = 0 for every b in [k, n)
    const bool pxi = par64(s.Gb & xi);
    u64 vxi = ballot(act && pxi);
    const bool in_span = ballot(!act && x < s.n && pxi) == 0;
    u64 vze = ballot(act && par64(s.G & zeta));
    bool inx = (vxi >> x) & 1;
    int w = 2 * m + 4 * par64(zeta & s.h);
    w += 2 * __popcll(ballot(inx && (s.d & 2))) + 4 * __popcll(ballot(inx && (s.d & 4)));
    w += 4 * __popcll(ballot(inx && par64(s.J & vxi & lowmask(x))));
    w &= 7;
    u64 eta = ballot(x < s.n && par64(s.J & vxi)) ^ vze;
    if (in_span) {
        if (w == 0 || w == 4) {
            u64 gamma = wave_xor((act && ((eta >> x) & 1)) ? s.Gb : 0ull);
            int alpha = ((w >> 2) + __popcll(gamma & s.h)) & 1;
            return rs_shrink(s, gamma, alpha, false);
        }
        int sigma = 2 - w / 2;
        s.Q = (((s.Q + sigma) % 8) + 8) % 8;
        if (act && ((eta >> x) & 1)) s.d = (s.d - 2 * sigma + 8) & 7;
        if ((eta >> x) & 1) s.J ^= eta;
        return 2;
    }
    rs_extend(s, xi);
    const int last = s.k - 1;
    int newD = (2 * m + 4 * par64(zeta & xi) + 4 * par64(zeta & s.h)) & 7;
    if (x == last) { s.d = newD; s.J = vze; }
    if (x < s.n) s.J = (s.J & ~bit(last)) | (((vze >> x) & 1ull) << last);
    if (x == last) s.J = (s.J & ~bit(last)) | ((u64)(m & 1) << last);
    return 2;
}

// stabilizer.c:689-756 with a Philox stream (same draw order as the oracle):
// x (53 bits) -> d; lazy shrinks by random xi until dim n-d (cap 4096 draws);
// h, D1, D2; J lower rows i < k; J_ii = [D_i in {2,6}]; Q = 0.
// Returns false if the cap was hit.
DEV bool rs_random_state(RState &s, int n, const uint64_t *thr, uint64_t key, uint64_t stream) {
    const int x = lane_id();
    Stream rng;
    rng.init(key, stream, 1u);
    const u64 nm = lowmask(n);
    u64 r = rng.next64() >> 11;
    int d = 0;
    while (d < n && !(r + 1 <= thr[d])) d++;
    const int k = n - d;
    rs_init(s, n);
    int iters = 0;
    while (s.k > k) {
        if (++iters > 4096) return false;
        u64 xi = rng.next64() & nm;
        rs_shrink(s, xi, 0, true);
    }
    s.h = rng.next64() & nm;
    u64 D1 = rng.next64() & nm;
    u64 D2 = rng.next64() & nm;
    s.d = 2 * (int)((D1 >> x) & 1ull) + 4 * (int)((D2 >> x) & 1ull);
    for (int i = 0; i < k; i++) {
        u64 lo = rng.next64() & lowmask(i);
        if (x == i) s.J |= lo;
        if ((lo >> x) & 1ull) s.J |= bit(i);
    }
    if (x < k && ((D1 >> x) & 1ull)) s.J |= bit(x);
    s.Q = 0;
    return true;
}

// ---------------------------------------------------- active-mask layout
// Per lane x: G row x and J row x (off-diagonal bits only are meaningful: J
// stays symmetric, so bit x of row i is J_xi) and D_x in Z8 (always even).
// Uniform: the active rows, h and Q.
//
// Packed form (PK = true, n <= 60): D_x lives in bits 61..63 of the J row,
// so one broadcast of row a gives both the J-row and D_a, and adding to D is
// a 64-bit add on the row (mod 8 in the top three bits, no carry into the
// row).  Unpacked form (PK = false): D_x in its own register.
#define CSE_ROWMASK 0x1FFFFFFFFFFFFFFFull

struct AState {
    u64 G, J;
    int d;      // PK = false only
    u64 act;
    u64 h;
    int Q;
};

// per-lane predicate "lane in mask" for a wave-uniform mask (no compare)
DEV bool lane_in(u64 mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

// all-ones if b else 0 (uniform integer select)
DEV u64 allmask(int b) { return 0ull - (u64)(b & 1); }

template <bool PK> DEV int d_of_row(u64 row, const AState &s, int a) {
    if (PK) return (int)(row >> 61);
    return rl32(s.d, a);
}
template <bool PK> DEV u64 row_bits(u64 row) { return PK ? (row & CSE_ROWMASK) : row; }
// D_x += v on lanes in U (v uniform, 0..7)
template <bool PK> DEV void d_add(AState &s, u64 U, int v) {
    if (PK) {
        if (lane_in(U)) s.J += (u64)v << 61;
    } else {
        if (lane_in(U)) s.d = (s.d + v) & 7;
    }
}
// D_x += 4 on lanes in U
template <bool PK> DEV void d_add4(AState &s, u64 U) {
    if (PK) {
        if (lane_in(U)) s.J ^= bit(63);
    } else {
        if (lane_in(U)) s.d ^= 4;
    }
}

// Basis change R = I + sum_{b in Sp} e_b e_i^T of the quadratic form
// (updateDJ of the reference, stabilizer.c:129-160, for a rank-1 R):
//   D_b += D_i + 4 J_bi, J <- R J R^T  (and G_b ^= G_i when withG).
// J_ii is taken as [D_i in {2,6}].  Returns M = column i of R J (bit x =
// (RJ)_xi), which fixing z_i needs; *di_out = D_i.
template <bool PK> DEV u64 as_basis_change(AState &s, int i, u64 Sp, bool withG, int *di_out) {
    const u64 bi = bit(i);
    const u64 Jr = rl64(s.J, i);
    const int di = d_of_row<PK>(Jr, s, i);
    const int jii = (di >> 1) & 1;
    const u64 Jc = (row_bits<PK>(Jr) & ~bi) | ((u64)jii << i);
    const u64 M = Jc ^ (Sp & allmask(jii));
    if (withG) {
        const u64 Gi = rl64(s.G, i);
        if (lane_in(Sp)) s.G ^= Gi;
    }
    // rows b in Sp: J_b ^= Jc, D_b += D_i + 4 J_bi (J_bi = bit b of Jc)
    if (PK) {
        if (lane_in(Sp)) s.J = (s.J ^ Jc) + ((u64)di << 61);
        if (lane_in(Sp & Jc)) s.J ^= bit(63);
    } else {
        const int four = lane_in(Jc) ? 4 : 0;
        if (lane_in(Sp)) {
            s.d = (s.d + di + four) & 7;
            s.J ^= Jc;
        }
    }
    // columns: rows x with (RJ)_xi = 1 get J_x ^= Sp
    if (lane_in(M)) s.J ^= Sp;
    *di_out = di;
    return M;
}

// Basis change as above, then fix z_i = c and retire row i.
template <bool PK> DEV void as_eliminate(AState &s, int i, u64 Sp, int c, bool withG) {
    int di;
    const u64 M = as_basis_change<PK>(s, i, Sp, withG, &di);
    if (c) {
        s.Q = (s.Q + di) & 7;
        d_add4<PK>(s, M);
        if (withG) s.h ^= rl64(s.G, i);
    }
    s.act &= ~bit(i);
}

// Restrict to {x : x . xi = alpha}.  Returns false if the restriction is empty.
template <bool PK> DEV bool as_restrict(AState &s, u64 xi, int alpha) {
    u64 S = ballot(lane_in(s.act) && par64(s.G & xi));
    int beta = (alpha ^ __popcll(s.h & xi)) & 1;
    if (S == 0) return beta == 0;
    int i = 63 - __clzll(S);
    as_eliminate<PK>(s, i, S & ~bit(i), beta, true);
    return true;
}

// Adds 4 x_u x_v (as a function of the affine coordinates) to the form:
// the coupling of an |00>+|01>+|10>-|11> pair (stateprep.c:45-50).
template <bool PK> DEV void as_add_coupling(AState &s, int u, int v) {
    const int gu = (int)((s.G >> u) & 1ull), gv = (int)((s.G >> v) & 1ull);
    const u64 cu = ballot(lane_in(s.act) && gu), cv = ballot(lane_in(s.act) && gv);
    const int hu = (int)((s.h >> u) & 1ull), hv = (int)((s.h >> v) & 1ull);
    s.Q = (s.Q + 4 * (hu & hv)) & 7;
    // linear part: 4 z_a (hu g_va + hv g_ua + g_ua g_va)
    const u64 lin = (cv & allmask(hu)) ^ (cu & allmask(hv)) ^ (cu & cv);
    d_add4<PK>(s, lin);
    // quadratic part: J_ab ^= g_ua g_vb + g_ub g_va
    if (lane_in(cu)) s.J ^= cv;
    if (lane_in(cv)) s.J ^= cu;
}

// sum_{z in F2^act} w^{q(z)} by eliminating one variable at a time:
//  D_a in {2,6}: sum over z_a gives sqrt2 w^{s(1-2l)}, l = parity of the
//    J-row; expanding the parity to second order updates D and J (rank 1);
//  D_a in {0,4}: sum over z_a gives 2 [l = D_a/4]; a non-empty J-row makes
//    that a linear constraint, solved by as_eliminate on its last member.
// Returns eps; p is the exponent of sqrt2, m the phase (mod 8, includes Q).
// Packed layout, branch-free on the lanes: basis change of pivot i over Sp
// (as as_basis_change), then z_i = c.  Rows b in Sp: J_b ^= Jc,
// D_b += D_i + 4 J_bi (and G_b ^= G_i when WITHG); rows x in M: J_x ^= Sp;
// c = 1 adds D_i to Q, 4 to D_x on M (and G_i to h when WITHG).  The caller
// retires row i.
template <bool WITHG> DEV void pk_eliminate(AState &s, int i, u64 Sp, unsigned c) {
    const u64 Jr = rl64(s.J, i);
    const unsigned di = (unsigned)(Jr >> 61);
    const u64 jii = (u64)((di >> 1) & 1u);
    const u64 Jc = (Jr & CSE_ROWMASK & ~bit(i)) | (jii << i);
    const u64 M = Jc ^ (Sp & (0ull - jii));
    const bool inS = lane_in(Sp), inSJ = lane_in(Sp & Jc), inM = lane_in(M);
    const unsigned jlo = (unsigned)s.J, jhi = (unsigned)(s.J >> 32);
    const u64 X = (inS ? Jc : 0ull) ^ (inM ? Sp : 0ull);
    unsigned nhi = (jhi ^ (unsigned)(X >> 32)) + (inS ? (di << 29) : 0u);
    nhi ^= (inSJ ? 0x80000000u : 0u) ^ ((inM && c) ? 0x80000000u : 0u);
    s.J = ((u64)nhi << 32) | (jlo ^ (unsigned)X);
    s.Q = (s.Q + (int)(di & (0u - c))) & 7;
    if (WITHG) {
        const u64 Gi = rl64(s.G, i);
        s.G = inS ? (s.G ^ Gi) : s.G;
        s.h ^= Gi & (0ull - (u64)c);
    }
}

// as_restrict on the packed layout for xi = e_q: {x : x_q = alpha}.  The
// rows in B (variables fixed to 1 and retired beforehand, never modified
// since) still contribute to x_q: their part of h is the parity of column q
// over B, taken from the same ballot as the pivot set.
DEV bool pk_restrict_coord(AState &s, int q, unsigned alpha, u64 B) {
    const u64 col = ballot((s.G >> q) & 1ull);
    const u64 S = col & s.act;
    const unsigned beta = (alpha ^ (unsigned)(s.h >> q) ^ (unsigned)__popcll(col & B)) & 1u;
    if (!S) return beta == 0;
    const int i = __builtin_ctzll(S);
    pk_eliminate<true>(s, i, S & ~bit(i), beta);
    s.act &= ~bit(i);
    return true;
}

//
// Packed layout (D_x in bits 61..63 of J_x), the elimination loop scheduled
// by hand: the compiler's lowering of this loop spends most of its scalar
// instructions on control-flow bookkeeping, and scalar issue bounds the
// sweep.  The loop takes, per step, the lowest active variable a:
//  D_a in {2,6} (bit 62 of J_a): factor sqrt2 w^{s}, s = +1 for 2, -1 for 6;
//    rows b in eta = J_a & act get J_b ^= eta and D_b -= D_a;
//  D_a in {0,4}, eta != 0: factor 2 and the constraint eta . z = D_a/4,
//    solved for i = min(eta) with Sp = eta \ {i} (pk_eliminate, c = D_a/4);
//  D_a in {0,4}, eta = 0: factor 2 (D_a = 0) or the sum vanishes (D_a = 4).
// Column i and row i are dead after a step, so Jc keeps the stored diagonal
// bit of row i, M may contain i, and eta (which contains i) stands for Sp.
// Variable 60 is the end row: it is always active, it is the highest, and
// its row (J = 0, D = 4) sends the loop to the exit through the vanishing
// branch, so no step tests for an empty mask (n <= 60).  The exit leaves its
// index in s56: 60 at the end row, the vanishing variable's index otherwise.
// act also keeps bit 63, so bit 63 of eta = J_a & act is D_a's top bit: in a
// {2,6} step, (J_b ^ eta) - 2 (mod 8 in the D field) is D_b - D_a for both
// values (D ^ 4 - 2 = D - 6), and in a constrained step it is c, so eta's
// bit 63 adds 4 c to D on the rows of M.
//
// With nA2, nA6 the {2,6} steps with D_a = 2 and 6, nC the steps without a
// constraint and n0 variables, the result is 2^{(n0 + nC)/2} w^m with
// m = Q + nA2 - nA6 + sum_B c D_i over the constrained steps.  W (s60) is
// the term being built: bits 29..31 hold Q + nA2 + nA6 + sum_B c D_i (each
// {2,6} step adds 1 there and its D_a top bit as the carry into bit 0, so
// m is bits 29..31 - 2 nA6), bits 7..27 hold base - n0 - nC in units of
// 128, and bits 0..6 hold 1 + nA6 (nA6 <= 60).  Register use: s56 pivot index, then D_i << 29,
// s[58:59] act, s[62:63] J_a, s[64:65] eta, s[66:67] Jc, v40 = -2 << 29,
// v41 = 1 << 31.  A
// VALU instruction reads an SGPR that v_readlane wrote only two or more
// instructions later.  Every step clears a bit of act below 60.
// CSE_ELIM_LOOP_BODY(S, CNEXT) is the loop with its labels suffixed by S
// (two copies can share an asm block) and CNEXT the instruction that
// follows a step without a constraint (the branch back to the top).
#define CSE_END_ROW 60
// The loop's head: the lowest active variable a leaves act, its row J_a
// (D_a in the top bits) is read, eta = J_a & act, SCC = D_a in {2,6}.  Each
// step ends with its own copy of the head, so a step goes on to the next
// one with one branch (to L_A, or on to L_nonA) instead of a jump back to a
// shared head and a second dispatch.
#define CSE_ELIM_HEAD                                                       \
    "s_ff1_i32_b64 s56, s[58:59]\n\t"                                       \
    "s_bitset0_b64 s[58:59], s56\n\t"                                       \
    "v_readlane_b32 s62, %[jlo], s56\n\t"                                   \
    "v_readlane_b32 s63, %[jhi], s56\n\t"                                   \
    "s_and_b64 s[64:65], s[62:63], s[58:59]\n\t"                            \
    "s_bitcmp1_b32 s63, 30\n\t"
#define CSE_ELIM_NEXT(S)                                                    \
    CSE_ELIM_HEAD                                                           \
    "s_cbranch_scc1 L_A" S "%=\n\t"                                          \
    "s_branch L_nonA" S "%=\n"
#define CSE_ELIM_LOOP_BODY(S, CNEXT)                                        \
    "L_top" S "%=:\n\t"                                                     \
    CSE_ELIM_HEAD                                                           \
    "s_cbranch_scc0 L_nonA" S "%=\n"                                        \
    "L_A" S "%=:\n\t"                                                       \
    /* D_a in {2,6}: rows in eta take J ^= eta, D -= D_a; m += 1, and       \
       the carry in (SCC = D_a's top bit) counts D_a = 6 in W's low bits */ \
    "v_xor_b32 %[vt0], s64, %[jlo]\n\t"                                     \
    "v_xad_u32 %[vt1], %[jhi], s65, v40\n\t"                                \
    "v_cndmask_b32_e64 %[jlo], %[jlo], %[vt0], s[64:65]\n\t"                \
    "v_cndmask_b32_e64 %[jhi], %[jhi], %[vt1], s[64:65]\n\t"                \
    "s_bitcmp1_b32 s63, 31\n\t"                                             \
    "s_addc_u32 s60, s60, 0x20000000\n\t"                                   \
    CSE_ELIM_HEAD                                                           \
    "s_cbranch_scc1 L_A" S "%=\n"                                           \
    "L_nonA" S "%=:\n\t"                                                    \
    /* eta without bit 63: a constraint? */                                 \
    "s_lshl_b64 s[66:67], s[64:65], 1\n\t"                                  \
    "s_cbranch_scc0 L_C" S "%=\n\t"                                         \
    /* D_a in {0,4} with a constraint: pivot i = min(eta).  Row i and       \
       column i are dead from here on, so eta itself serves as Sp;          \
       its bit 63 is c = D_a / 4, so D += 4 c on the rows of M comes        \
       with the XOR of Sp.  bitop3 0x78 is S0 ^ (S1 & S2), 0x60 is          \
       S0 & (S1 ^ S2). */                                                   \
    "s_ff1_i32_b64 s56, s[64:65]\n\t"                                       \
    "s_bitset0_b64 s[58:59], s56\n\t"                                       \
    "v_readlane_b32 s66, %[jlo], s56\n\t"                                   \
    "v_readlane_b32 s67, %[jhi], s56\n\t"                                   \
    "v_cndmask_b32_e64 %[vt0], 0, -1, s[64:65]\n\t"                         \
    "s_and_b32 s56, s67, 0xe0000000\n\t"                                    \
    "s_bitcmp1_b32 s67, 30\n\t"                                             \
    "s_cbranch_scc1 L_jii" S "%=\n\t"                                       \
    /* D_i in {0,4}: M = Jc, and D += D_i on Sp is bit 63 of Jc. Rows       \
       of M take J ^= Sp with bit 63 flipped where x is in Sp too           \
       (the 4 J_xi of the rows in Sp). */                                   \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t"                         \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t"              \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t"              \
    "v_bitop3_b32 %[vt1], s65, %[vt0], v41 bitop3:0x78\n\t"                 \
    "s_bitcmp1_b32 s63, 31\n\t"                                             \
    "s_cbranch_scc0 L_noc0" S "%=\n\t"                                      \
    "s_add_u32 s60, s60, s56\n"                                             \
    "L_noc0" S "%=:\n\t"                                                    \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t"              \
    "v_bitop3_b32 %[jhi], %[jhi], %[vt1], %[vt3] bitop3:0x78\n\t"           \
    CSE_ELIM_NEXT(S)                                                        \
    /* D_i in {2,6}: M = Jc ^ Sp; rows in Sp take J ^= Jc and               \
       D += D_i + 4 J_xi */                                                 \
    "L_jii" S "%=:\n\t"                                                     \
    "s_and_b32 s67, s67, 0x1fffffff\n\t"                                    \
    "v_cndmask_b32_e64 %[vt3], 0, -1, s[66:67]\n\t"                         \
    "v_bitop3_b32 %[jlo], %[jlo], s66, %[vt0] bitop3:0x78\n\t"              \
    "v_bitop3_b32 %[jhi], %[jhi], s67, %[vt0] bitop3:0x78\n\t"              \
    "v_and_b32 %[vt1], 0x80000000, %[vt3]\n\t"                              \
    "v_bitop3_b32 %[vt2], %[vt0], s56, %[vt1] bitop3:0x60\n\t"              \
    "v_add_u32 %[jhi], %[jhi], %[vt2]\n\t"                                  \
    "v_xor_b32 %[vt3], %[vt3], %[vt0]\n\t"                                  \
    "s_bitcmp1_b32 s63, 31\n\t"                                             \
    "s_cbranch_scc0 L_noc1" S "%=\n\t"                                      \
    "s_add_u32 s60, s60, s56\n"                                             \
    "L_noc1" S "%=:\n\t"                                                    \
    "v_bitop3_b32 %[jlo], %[jlo], s64, %[vt3] bitop3:0x78\n\t"              \
    "v_bitop3_b32 %[jhi], %[jhi], s65, %[vt3] bitop3:0x78\n\t"              \
    CSE_ELIM_NEXT(S)                                                        \
    /* D_a in {0,4} without a constraint: the sum vanishes (D_a = 4,        \
       or the end row) or nC += 1 */                                        \
    "L_C" S "%=:\n\t"                                                       \
    "s_bitcmp1_b32 s63, 31\n\t"                                             \
    "s_cbranch_scc1 L_end%=\n\t"                                            \
    "s_add_u32 s60, s60, 0xffffff80\n\t"                                    \
    CNEXT
#define CSE_ELIM_LOOP CSE_ELIM_LOOP_BODY("", "s_branch L_top%=\n")

// The loop on an AState (records of n <= 60, row 60 the end row); p and m
// as as_expsum.  W's p field counts down from the base 128 (p = n0 + nC
// <= 2 n0 <= 120), so p = 128 - (bits 7..27 of W).
DEV int as_expsum_packed(AState &s, int &p, int &m) {
    const u64 act0 = uniform64(s.act);
    u64 act = act0 | bit(CSE_END_ROW) | bit(63);
    s.act = 0;
    const int n0 = __popcll(act0);
    unsigned jlo = (unsigned)s.J, jhi = (unsigned)(s.J >> 32);
    unsigned W = (unsigned)__builtin_amdgcn_readfirstlane(
        (int)(((unsigned)s.Q << 29) + 128u * (unsigned)(128 - n0) + 1u));
    unsigned vt0, vt1, vt2, vt3;
    int endv;
    asm volatile(
        "v_mov_b32 v40, 0xc0000000\n\t"
        "v_mov_b32 v41, 0x80000000\n"
        CSE_ELIM_LOOP
        "L_end%=:\n\t"
        "s_mov_b32 %[endv], s56\n\t"
        : "+{s[58:59]}"(act), "+{s60}"(W), [jlo] "+v"(jlo), [jhi] "+v"(jhi), [vt0] "=&v"(vt0),
          [vt1] "=&v"(vt1), [vt2] "=&v"(vt2), [vt3] "=&v"(vt3), [endv] "=&s"(endv)
        :
        : "s56", "s62", "s63", "s64", "s65", "s66", "s67", "v40", "v41", "scc");
    if (endv != CSE_END_ROW) {
        p = 0;
        m = 0;
        return 0;
    }
    p = 128 - (int)((W & 0x0fffffffu) >> 7);
    m = (int)((W >> 29) - 2u * ((W & 0x7fu) - 1u)) & 7;
    return 1;
}

template <bool PK> DEV int as_expsum(AState &s, int &p, int &m) {
    if (PK) return as_expsum_packed(s, p, m);
    int pp = 0, macc = 0, zero = 0;
    u64 act = s.act;
    while (act) {
        const int a = __ffsll((long long)act) - 1;
        act &= ~bit(a);
        const u64 Ja = rl64(s.J, a);
        const u64 eta = Ja & act;
        const int da = d_of_row<PK>(Ja, s, a);
        const int h4 = (da >> 2) & 1;
        if (da & 2) {
            // D_a = 2 (s = +1) or 6 (s = -1): factor sqrt2 w^s; D_b -= 2s
            // and J_bc ^= 1 on the J-row eta of a
            pp += 1;
            macc += h4 ? 7 : 1;
            const int delta = h4 ? 2 : 6;
            if (PK) {
                if (lane_in(eta)) s.J = (s.J ^ eta) + ((u64)delta << 61);
            } else if (lane_in(eta)) {
                s.d = (s.d + delta) & 7;
                s.J ^= eta;
            }
        } else if (eta) {
            // factor 2 and the constraint sum_{b in eta} z_b = D_a/4
            pp += 2;
            const int i = 63 - __clzll(eta);
            act &= ~bit(i);
            int di;
            const u64 M = as_basis_change<PK>(s, i, eta & ~bit(i), false, &di);
            s.Q = (s.Q + (di & -h4)) & 7;
            d_add4<PK>(s, M & allmask(h4));
        } else {
            // factor 1 + w^{D_a}: 2 if D_a = 0, the whole sum vanishes if 4
            pp += 2;
            zero |= h4;
            act &= (u64)h4 - 1ull;
        }
    }
    s.act = 0;
    p = zero ? 0 : pp;
    m = zero ? 0 : ((s.Q + macc) & 7);
    return zero ? 0 : 1;
}

}  // namespace cse
