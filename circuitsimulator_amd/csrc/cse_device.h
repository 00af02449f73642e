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
// XOR of v over all lanes (butterfly); returned as a wave-uniform value
DEV u64 wave_xor(u64 v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v ^= shfl_xor64(v, off);
    return uniform64(v);
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
    u64 vxi = ballot(act && par64(s.Gb & xi));
    u64 vze = ballot(act && par64(s.G & zeta));
    u64 xip = wave_xor(((vxi >> x) & 1) ? s.G : 0ull);
    bool inx = (vxi >> x) & 1;
    int w = 2 * m + 4 * par64(zeta & s.h);
    w += 2 * __popcll(ballot(inx && (s.d & 2))) + 4 * __popcll(ballot(inx && (s.d & 4)));
    w += 4 * __popcll(ballot(inx && par64(s.J & vxi & lowmask(x))));
    w &= 7;
    u64 eta = ballot(x < s.n && par64(s.J & vxi)) ^ vze;
    if (xi == xip) {
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
struct AState {
    u64 G, J;   // per lane
    int d;      // per lane, Z8
    u64 act;    // uniform: active rows
    u64 h;      // uniform
    int Q;      // uniform
};

// Basis change R = I + sum_{b in Sp} e_b e_i^T (updateDJ of the reference,
// stabilizer.c:129-160, for a rank-1 R; G_b ^= G_i when withG).  The J
// diagonal is not trusted: J_ii is taken from D_i.  Returns (through rjxi)
// bit i of each row of R J, which the fixation needs.
DEV int as_basis_change(AState &s, int i, u64 Sp, bool withG) {
    const int x = lane_id();
    const bool inSp = (Sp >> x) & 1;
    u64 Ji = rl64(s.J, i);
    const int di = rl32(s.d, i);
    const u64 jii = (u64)((di >> 1) & 1);
    Ji = (Ji & ~bit(i)) | (jii << i);
    u64 Gi = 0;
    if (withG) Gi = rl64(s.G, i);
    const int jxi = (int)((s.J >> i) & 1ull);
    if (inSp) {
        s.d = (s.d + di + 4 * jxi) & 7;
        s.J ^= Ji;
        if (withG) s.G ^= Gi;
    }
    const int rjxi = (int)((s.J >> i) & 1ull);
    if (rjxi) s.J ^= Sp;
    return rjxi;
}

// Basis change as above, then fix z_i = c and retire row i.
DEV void as_eliminate(AState &s, int i, u64 Sp, int c, bool withG) {
    const int di = rl32(s.d, i);
    const int rjxi = as_basis_change(s, i, Sp, withG);
    if (c) {
        s.Q = (s.Q + di) & 7;
        s.d = (s.d + 4 * rjxi) & 7;
        if (withG) s.h ^= rl64(s.G, i);
    }
    s.act &= ~bit(i);
}

// Restrict to {x : x . xi = alpha}.  Returns false if the restriction is empty.
DEV bool as_restrict(AState &s, u64 xi, int alpha) {
    const int x = lane_id();
    u64 S = ballot(((s.act >> x) & 1) && par64(s.G & xi));
    int beta = (alpha ^ __popcll(s.h & xi)) & 1;
    if (S == 0) return beta == 0;
    int i = 63 - __clzll(S);
    as_eliminate(s, i, S & ~bit(i), beta, true);
    return true;
}

// Adds 4 x_u x_v (as a function of the affine coordinates) to the form:
// the coupling of an |00>+|01>+|10>-|11> pair (stateprep.c:45-50).
DEV void as_add_coupling(AState &s, int u, int v) {
    const int x = lane_id();
    const bool a = (s.act >> x) & 1;
    const int gu = (int)((s.G >> u) & 1ull), gv = (int)((s.G >> v) & 1ull);
    u64 cu = ballot(a && gu), cv = ballot(a && gv);
    const int hu = (int)((s.h >> u) & 1ull), hv = (int)((s.h >> v) & 1ull);
    s.Q = (s.Q + 4 * (hu & hv)) & 7;
    if (a) {
        int lin = (hu & gv) ^ (hv & gu) ^ (gu & gv);
        s.d = (s.d + 4 * lin) & 7;
        u64 t = (gu ? cv : 0ull) ^ (gv ? cu : 0ull);
        s.J ^= t;
    }
}

// sum_{z in F2^act} w^{q(z)} by eliminating one variable at a time:
//  D_a in {2,6}: sum over z_a gives sqrt2 w^{s(1-2l)}, l = parity of the
//    J-row; expanding the parity to second order updates D and J (rank 1);
//  D_a in {0,4}: sum over z_a gives 2 [l = D_a/4]; a non-empty J-row makes
//    that a linear constraint, solved by as_eliminate on its last member.
// Returns eps; p is the exponent of sqrt2, m the phase (mod 8, includes Q).
DEV int as_expsum(AState &s, int &p, int &m) {
    const int x = lane_id();
    int pp = 0, macc = 0;
    while (s.act) {
        const int a = __ffsll((long long)s.act) - 1;
        const u64 Ja = rl64(s.J, a);
        const int da = rl32(s.d, a);
        s.act &= ~bit(a);
        const u64 eta = Ja & s.act;
        if (da & 2) {
            const int sg = (da == 2) ? 1 : 7;  // +1 or -1 (mod 8)
            pp += 1;
            macc += sg;
            if ((eta >> x) & 1) {
                s.d = (s.d - 2 * ((da == 2) ? 1 : -1) + 8) & 7;
                s.J ^= eta;
            }
        } else {
            pp += 2;
            const int c = da >> 2;
            if (eta == 0) {
                if (c) { p = 0; m = 0; return 0; }
            } else {
                const int i = 63 - __clzll(eta);
                as_eliminate(s, i, eta & ~bit(i), c, false);
            }
        }
    }
    p = pp;
    m = (s.Q + macc) & 7;
    return 1;
}

}  // namespace cse
