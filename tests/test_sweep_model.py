"""CPU model of the sweep's record path, checked against the oracle.

The sweep kernel (circuitsimulator_amd/csrc/cse_kernels.hip, cse_device.h)
computes <phi_j|theta'> from an elimination-ready record of theta': the basis
of K in reduced row echelon form with row a stored at its pivot column, the
quadratic form on it, and D packed into bits 61..63 of the J rows.  For
phi_j = prepL(j) the pivot coordinates outside supp(x_j) are fixed in one
batch, the other zero coordinates are restrictions, and the remaining
Gaussian sum is eliminated one variable at a time by a hand-scheduled loop
that ends at row 60 (J = 0, D = 4).

This file restates that path in Python register for register: phiL_batch
(the asm block of records with at most two non-pivot columns: per-lane
tables of x_j and the term-independent part of the W word m * 2^29 +
128 np + 1, fixation with one edge ballot, restrictions as constrained steps
with the second column updated in scalar registers, the loop, W parked and
unpacked later), its exact-mode form (every non-pivot column outside
supp(x_j) a restriction, columns from working G rows, the batch of empty
first restrictions, the loop's batch of uncoupled variables), the |H^t>
block phiH_batch (pair constraints and couplings), the fallback of records
with more non-pivot columns (restrictions in C++, then the same loop), and
the exact-mode record built directly in echelon form (rs_project_rref),
and the wave's 64-ary search for the instance of a record or item
(find_instance).  Every term is checked against the oracle's innerProduct (libcirc/
stabilizer/stabilizer.c:589-667), every directly built record against the
oracle's projection, and the step kinds are counted so that every branch is
known to be exercised.
"""
import collections

import numpy as np
import pytest

import orc

MASK61 = (1 << 61) - 1
U32 = 0xFFFFFFFF


def bits(m):
    while m:
        b = (m & -m).bit_length() - 1
        yield b
        m &= m - 1


def low(m):
    return (m & -m).bit_length() - 1


def popc(m):
    return bin(m).count("1")


class Form:
    """Active-mask layout: rows G, J (symmetric), D per variable, act, h, Q."""

    def __init__(self):
        self.G = [0] * 64
        self.J = [0] * 64
        self.d = [0] * 64
        self.act = 0
        self.h = 0
        self.Q = 0
        self.k = 0

    def copy(self):
        c = Form()
        c.G, c.J, c.d = self.G[:], self.J[:], self.d[:]
        c.act, c.h, c.Q, c.k = self.act, self.h, self.Q, self.k
        return c


def basis_change(s, i, Sp, withG):
    """R = I + sum_{b in Sp} e_b e_i^T (updateDJ for a rank-1 R,
    stabilizer.c:129-160); returns M = column i of R J and D_i."""
    Jr, di = s.J[i], s.d[i]
    jii = (di >> 1) & 1
    Jc = (Jr & ~(1 << i)) | (jii << i)
    M = Jc ^ (Sp if jii else 0)
    if withG:
        for b in bits(Sp):
            s.G[b] ^= s.G[i]
    for b in bits(Sp):
        s.d[b] = (s.d[b] + di + 4 * ((Jc >> b) & 1)) & 7
        s.J[b] ^= Jc
    for x in bits(M):
        s.J[x] ^= Sp
    return M, di


def to_record(th):
    """The device's to_record: RREF by rank-1 basis changes, rows re-indexed
    by pivot column."""
    s = Form()
    k = th.k
    for a in range(64):
        s.G[a], s.J[a] = th.G[a], th.J[a]
        s.d[a] = th.D(a) if a < th.n else 0
    s.h, s.Q, s.act = th.h, th.Q, (1 << k) - 1
    piv, Pm = [0] * k, 0
    for a in range(k):
        pc = low(s.G[a])
        piv[a] = pc
        Pm |= 1 << pc
        Sp = sum(1 << x for x in range(k) if x != a and (s.G[x] >> pc) & 1)
        basis_change(s, a, Sp, True)
    r = Form()
    r.h, r.Q, r.act, r.k = s.h, s.Q & 7, Pm, k
    for a in range(k):
        r.G[piv[a]] = s.G[a]
        r.d[piv[a]] = s.d[a] & 6
        r.J[piv[a]] = sum(1 << piv[b] for b in range(k) if (s.J[a] >> b) & 1)
    return r


END = 60  # the end row of the elimination loop
TOP = 63  # kept in act: bit 63 of eta is D_a's top bit
U64 = (1 << 64) - 1
LIVE = (1 << END) - 1  # the variables' bits (the end row and bit 63 excluded)

# Algorithmic work of the fixation/elimination algorithm, in 64-bit word
# operations of a dense-row formulation (one operation per live row and per
# row-level operation of a step), counted under count["w_*"]:
#   fixation of the pivot coordinates outside supp(x_j): per record row, the
#     AND with the fixed set, its parity and the D update: 3 per row;
#   restriction on a non-pivot column: the column's bits over the live rows
#     (2 per row), then, unless it is empty, a constrained elimination (5 per
#     live row, as a B step);
#   A step (D_a in {2,6}): the pivot, then per live row the membership in
#     eta, J ^= eta and D -= 2: 1 + 3 per live row;
#   B step (D_a in {0,4} with a constraint): the pivot, then per live row the
#     membership in Sp, J ^= J_i, the D update, the membership in M and
#     J ^= Sp: 1 + 5 per live row;
#   C step (no constraint) and the loop's end: 1;
#   the batch of uncoupled variables (exact mode): one pass, 2 per live row.
W_STEP = {"A": 3, "B": 5}


def iso_batch(J, act, W, count):
    """The batch of CSE_ELIM_LOOP_ISO at a term's first C step: the active
    variables whose row meets no active variable below bit 61 (its own bit
    included) and whose bit 62 is clear, the end row and bit 63 excepted,
    leave act together with W -= 128 each; one of them with D = 4 (bit 63)
    ends the loop through the vanishing exit (s56 = 0).  Returns (act, W,
    vanished)."""
    hmask = ((act >> 32) & 0x1FFFFFFF) | 0x40000000
    Z = 0
    for x in range(64):
        if ((J[x] & act & U32) | ((J[x] >> 32) & hmask)) == 0:
            Z |= 1 << x
    Z &= act & ~((1 << END) | (1 << TOP))
    count["iso_batch"] += 1
    if any((J[x] >> 63) & 1 for x in bits(Z)):
        count["iso_vanish"] += 1
        return act, W, True
    count["iso_vars"] += popc(Z)
    return act & ~Z, (W - 128 * popc(Z)) & U32, False


def elim_loop(J, act, W, count, iso=False):
    """The hand-scheduled elimination loop (CSE_ELIM_LOOP) on packed rows
    J_x | D_x << 61, act holding the end row and bit 63; returns (s56, W):
    s56 is the index of the variable that ended the loop (END at the end
    row).  iso: CSE_ELIM_LOOP_ISO, the batch of uncoupled variables at the
    first C step."""
    while True:
        a = low(act)
        live = popc(act & LIVE)
        act &= ~(1 << a)
        Ja = J[a]
        hi = Ja >> 32
        eta = Ja & act                           # bit 63: D_a >> 2
        if (hi >> 30) & 1:                       # D_a in {2,6}
            count["A"] += 1
            count["w_A"] += 1 + W_STEP["A"] * live
            for x in bits(eta):
                lo = (J[x] & U32) ^ (eta & U32)
                h2 = (((J[x] >> 32) ^ (eta >> 32)) + 0xC0000000) & U32
                J[x] = (h2 << 32) | lo
            # s_addc: m += 1, and D_a's top bit (SCC) counted in bits 0..6
            W = (W + 0x20000000 + ((hi >> 31) & 1)) & U32
            continue
        if not (eta << 1) & U64:                 # no constraint
            if (hi >> 31) & 1:                   # D_a = 4, or the end row
                if a != END:
                    count["vanish"] += 1
                count["w_C"] += 1
                return a, W
            count["C"] += 1
            count["w_C"] += 1
            W = (W + 0xFFFFFF80) & U32
            if iso:
                iso = False
                count["w_iso"] += 2 * popc(act & LIVE)
                act, W, gone = iso_batch(J, act, W, count)
                if gone:
                    return 0, W
            continue
        count["B"] += 1                          # constraint eta . z = D_a/4
        count["w_B"] += 1 + W_STEP["B"] * live
        act &= ~(1 << low(eta))
        i = low(eta)
        Jr = J[i]
        s66, s67 = Jr & U32, Jr >> 32
        s56 = s67 & 0xE0000000
        s64, s65 = eta & U32, eta >> 32          # s65 bit 31: c
        c = (hi >> 31) & 1
        sp = [(eta >> x) & 1 for x in range(64)]
        if not (s67 >> 30) & 1:                  # D_i in {0,4}: bit 63 of J_i adds it
            count["B0"] += 1
            jc = (s67 << 32) | s66
            for x in range(64):
                if sp[x]:
                    J[x] ^= jc
            if c:
                W = (W + s56) & U32
            for x in range(64):
                if (jc >> x) & 1:
                    vt1 = s65 ^ (0x80000000 if sp[x] else 0)
                    J[x] ^= (vt1 << 32) | s64
        else:                                    # D_i in {2,6}
            count["B1"] += 1
            s67 &= 0x1FFFFFFF
            jc = (s67 << 32) | s66
            for x in range(64):
                lo, h2 = J[x] & U32, J[x] >> 32
                if sp[x]:
                    lo ^= s66
                    h2 ^= s67
                    h2 = (h2 + (s56 ^ (0x80000000 if (jc >> x) & 1 else 0))) & U32
                J[x] = (h2 << 32) | lo
            if c:
                W = (W + s56) & U32
            M = jc ^ eta
            for x in bits(M):
                J[x] ^= (s65 << 32) | s64


def restrict_step(J, S, c, W):
    """The restriction's elimination of i = min(S) over Sp = S with z_i = c
    (the asm's copy of the constrained step, S without bit 63): rows in Sp
    take J ^= Jc and D += D_i + 4 J_xi, rows in M take J ^= Sp (D += 4 when
    c), W += 128 (the variable leaves the exponent) + c D_i."""
    i = low(S)
    Jr = J[i]
    s66, s67 = Jr & U32, Jr >> 32
    s56 = s67 & 0xE0000000
    s67 &= 0x1FFFFFFF
    jc = (s67 << 32) | s66
    s64, s65 = S & U32, S >> 32
    for x in range(64):
        lo, h2 = J[x] & U32, J[x] >> 32
        if (S >> x) & 1:
            lo ^= s66
            h2 ^= s67
            h2 = (h2 + (s56 ^ (0x80000000 if (jc >> x) & 1 else 0))) & U32
        J[x] = (h2 << 32) | lo
    M = jc ^ (S if (s56 >> 30) & 1 else 0)
    W = (W + 0x80) & U32
    if c:
        s65 |= 0x80000000
        W = (W + s56) & U32
    for x in bits(M):
        J[x] ^= (s65 << 32) | s64
    return W


def unpack(W):
    """phiL_unpack: the packed term of a parked W word (0 stays 0): X from
    bits 7..27, m from bits 29..31 less twice the D_a = 6 count that bits
    0..6 hold above 1."""
    if not W:
        return 0
    m = ((W >> 29) - 2 * ((W & 0x7F) - 1)) & 7
    return ((((W & 0x0FFFFFFF) >> 7) << 4) + 1 + (m << 1)) & U32


def packed_rows(r):
    J = [(r.J[x] & MASK61) | ((r.d[x] & 7) << 61) for x in range(64)]
    J[END] = 4 << 61
    return J


def batch_consts(r, n):
    """The wave-uniform constants phiL_batch takes for a record with at most
    two non-pivot columns."""
    nonact = ((1 << n) - 1) & ~r.act
    qs = list(bits(nonact))
    col = [sum(1 << x for x in range(64) if (r.G[x] >> q) & 1) & r.act for q in qs] + [0, 0]
    return dict(J0=packed_rows(r), actR=r.act | (1 << END) | (1 << TOP), actH=r.act & r.h,
                D1=sum(1 << x for x in bits(r.act) if r.d[x] & 2),
                D2=sum(1 << x for x in bits(r.act) if r.d[x] & 4),
                nonact=nonact, col0=col[0], col1=col[1],
                hq=sum(((r.h >> q) & 1) << b for b, q in enumerate(qs)), Q=r.Q & 7, Kw=128 * r.k + 1)


def table_entry(C, xj):
    """phiL_tables for one lane: x_j with bits 60 and 63 set (so act =
    actR & x_j holds them; their counts in n0 and |x_j| cancel) and the
    term-independent part of W."""
    xv = xj | (1 << END) | (1 << TOP)
    n0 = popc(C["actR"] & xv)
    c1 = popc(C["D1"] & C["actH"] & ~xv)
    T = (((C["Q"] + 2 * c1) << 29) + 128 * (popc(xv) - n0) + C["Kw"]) & U32
    return xv, T


def batch_ip(C, xj, count):
    """One term of phiL_batch (cse_kernels.hip), register for register:
    fixation, up to two restrictions, the loop; returns the parked W word."""
    xj, T = table_entry(C, xj)
    J0 = C["J0"]
    B = C["actH"] & ~xj
    act = C["actR"] & xj
    count["w_fix"] += 3 * popc(C["actR"] & LIVE)
    J, E = [0] * 64, 0
    for x in range(64):
        jb = J0[x] & B
        E |= (popc(jb & ((1 << x) - 1)) & 1) << x
        J[x] = (J0[x] + ((popc(jb) & 1) << 63)) & U64
    W = (T + (popc((E ^ C["D2"]) & B) << 31)) & U32
    Zn = C["nonact"] & ~xj
    if Zn:
        q0 = low(C["nonact"])
        if (Zn >> q0) & 1:
            col, h, Zn = C["col0"], C["hq"] & 1, Zn & ~(1 << q0)
        else:
            col, h, Zn = C["col1"], C["hq"] >> 1, 0
        while True:
            count["restrict"] += 1
            S = col & act
            count["w_res"] += (2 + (5 if S else 0)) * popc(act & LIVE)
            beta = (popc(col & B) ^ h) & 1
            if S:
                i = low(S)
                act &= ~(1 << i)
                g = (C["col1"] >> i) & 1
                ncol, nh = C["col1"] ^ (S if g else 0), (C["hq"] >> 1) ^ (beta if g else 0)
                W = restrict_step(J, S, beta, W)
            else:
                if beta:
                    return 0
                ncol, nh = C["col1"], C["hq"] >> 1
            if not Zn:
                break
            count["restrict2"] += 1
            col, h, Zn = ncol, nh, 0
    s56, W = elim_loop(J, act, W, count)
    return W if s56 == END else 0


def batch_ip_gen(C, r, xj, count):
    """One term of phiL_batch<true> (the exact-mode sweep), register for
    register: fixation as in batch_ip, then every non-pivot column outside
    supp(x_j) in ascending order as a restriction whose column comes from the
    working G rows (reset from the record's at the first one) and whose h bit
    from the working h; an elimination of i over S also takes G_b ^= G_i on
    the rows of S (row i itself becomes 0; it is retired) and h ^= G_i when
    c = 1.  The loop is CSE_ELIM_LOOP_ISO.  Returns the parked W word."""
    xj, T = table_entry(C, xj)
    J0 = C["J0"]
    B = C["actH"] & ~xj
    act = C["actR"] & xj
    count["w_fix"] += 3 * popc(C["actR"] & LIVE)
    J, E = [0] * 64, 0
    for x in range(64):
        jb = J0[x] & B
        E |= (popc(jb & ((1 << x) - 1)) & 1) << x
        J[x] = (J0[x] + ((popc(jb) & 1) << 63)) & U64
    W = (T + (popc((E ^ C["D2"]) & B) << 31)) & U32
    pend = C["nonact"] & ~xj
    if pend:
        G, h = r.G[:], r.h
        first = True
        while pend:
            q = low(pend)
            pend &= pend - 1
            count["rg"] += 1
            col = sum(1 << x for x in range(64) if (G[x] >> q) & 1)
            S = col & act
            count["w_res"] += (2 + (5 if S else 0)) * popc(act & LIVE)
            c = (popc(col & B) ^ (h >> q)) & 1
            if not S:
                if c:
                    count["rg_vanish"] += 1
                    return 0
                count["rg_empty"] += 1
                if first:
                    # the first restriction is empty: every pending column
                    # that misses act, from the record's column table (lane q:
                    # the column's ballot over the pivot rows, h_q in bit 63)
                    count["rg_batch"] += 1
                    E = 0
                    for qq in bits(pend):
                        colt = (sum(1 << x for x in range(64) if (r.G[x] >> qq) & 1) & r.act) | (((r.h >> qq) & 1) << 63)
                        if colt & act & ((1 << 60) - 1) == 0:
                            E |= 1 << qq
                            if popc(colt & (B | (1 << 63))) & 1:
                                count["rg_batch_vanish"] += 1
                                return 0
                    count["rg_batch_cols"] += popc(E)
                    pend &= ~E
                first = False
                continue
            first = False
            i = low(S)
            act &= ~(1 << i)
            Gi = G[i]
            for x in bits(S):
                G[x] ^= Gi
            if c:
                count["rg_c"] += 1
                h ^= Gi
            if (J[i] >> 62) & 1:
                count["rg_jii"] += 1
            W = restrict_step(J, S, c, W)
            count["rg_elim"] += 1
    s56, W = elim_loop(J, act, W, count, iso=True)
    return W if s56 == END else 0


def fallback_ip(r, xj, n, count):
    """ip_phiL<true> with as_expsum_packed: the path of records with three or
    more non-pivot columns (and the layout the exact-mode terms use)."""
    F = r.act & ~xj
    B = F & r.h
    s = r.copy()
    count["w_fix"] += 3 * popc(r.act)
    edges = sum(1 for a in bits(B) for b in bits(B) if a < b and (r.J[a] >> b) & 1)
    s.Q = (r.Q + sum(r.d[a] for a in bits(B)) + 4 * edges) & 7
    for x in range(64):
        if popc(r.J[x] & B) & 1:
            s.d[x] ^= 4
    s.act = r.act & xj
    nonact = ((1 << n) - 1) & ~r.act
    for q in bits(nonact & ~xj):
        count["restrict"] += 1
        col = sum(1 << x for x in range(64) if (s.G[x] >> q) & 1)
        S = col & s.act
        count["w_res"] += (2 + (5 if S else 0)) * popc(s.act)
        beta = ((s.h >> q) ^ popc(col & B)) & 1
        if not S:
            if beta:
                return 0
            continue
        i = low(S)
        M, di = basis_change(s, i, S & ~(1 << i), True)
        if beta:
            s.Q = (s.Q + di) & 7
            for x in bits(M):
                s.d[x] ^= 4
            s.h ^= s.G[i]
        s.act &= ~(1 << i)
    n0 = popc(s.act)
    W = (((s.Q << 29) & U32) + 128 * (128 - n0) + 1) & U32
    s56, W = elim_loop(packed_rows(s), s.act | (1 << END) | (1 << TOP), W, count)
    if s56 != END:
        return 0
    p = 128 - ((W & 0x0FFFFFFF) >> 7)
    m = ((W >> 29) - 2 * ((W & 0x7F) - 1)) & 7
    return 1 + 2 * ((r.k + popc(xj) - p) * 8 + m)


def decode(v):
    if v == 0:
        return (0, 0, 0)
    v -= 1
    return (1, -(v >> 4), (v >> 1) & 7)


def xj_of(j, L, k):
    v = 0
    for rr in range(k):
        if (j >> (k - 1 - rr)) & 1:
            v ^= L[rr]
    return v


def run_model(t, r_gens, k, seed, nth, nj, only_fallback=False):
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r_gens)
    L = random_L(rng, t, k)
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    count = collections.Counter()
    checked = 0
    st = 0
    phis = [orc.prepL(j, t, L) for j in range(nj)]
    while checked < nth * nj:
        th = orc.random_state(t, seed, st, thr)
        st += 1
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        batch = popc(((1 << t) - 1) & ~rec.act) <= 2
        if batch and only_fallback:
            continue
        count["batch" if batch else "fallback"] += 1
        C = batch_consts(rec, t)
        for j in range(nj):
            xj = xj_of(j, L, k)
            e, p, m = orc.inner(th, phis[j])
            want = (e, p, m % 8) if e else (0, 0, 0)
            got = decode(fallback_ip(rec, xj, t, count))
            assert got == want, ("fallback", t, st - 1, j, got, want)
            if batch:
                got = decode(unpack(batch_ip(C, xj, count)))
                assert got == want, ("batch", t, st - 1, j, got, want)
            got = decode(unpack(batch_ip_gen(C, rec, xj, count)))
            assert got == want, ("general", t, st - 1, j, got, want)
            checked += 1
    return count


def run_exact_model(t, r_gens, k, seed, ni, nj):
    """Exact-mode records: theta'_i = Pi prepL(i) against prepL(j), the
    records with about t/2 non-pivot columns that the exact-mode sweep runs
    through the general restriction loop."""
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r_gens)
    L = random_L(rng, t, k)
    Po = orc.make_projector(ph, xs, zs, t)
    count = collections.Counter()
    chi = 1 << k
    js = sorted(set(int(v) for v in rng.integers(0, chi, size=nj)))
    phis = {j: orc.prepL(j, t, L) for j in js}
    for i in sorted(set(int(v) for v in rng.integers(0, chi, size=ni))):
        th = orc.prepL(i, t, L)
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        C = batch_consts(rec, t)
        count["records"] += 1
        for j in js:
            e, p, m = orc.inner(th, phis[j])
            want = (e, p, m % 8) if e else (0, 0, 0)
            xj = xj_of(j, L, k)
            got = decode(unpack(batch_ip_gen(C, rec, xj, count)))
            assert got == want, ("general", t, i, j, got, want)
            assert decode(fallback_ip(rec, xj, t, count)) == want, ("fallback", t, i, j)
    return count


@pytest.mark.parametrize("t,r,k,seed", [(48, 24, 12, 5), (20, 6, 8, 6), (9, 3, 4, 7), (60, 20, 10, 8)])
def test_record_path_matches_oracle(t, r, k, seed):
    count = run_model(t, r, k, seed, nth=3, nj=min(24, 1 << k))
    # every kind of step of the loop ran at least once over the cases
    assert count["A"] > 0 and count["B"] > 0


def test_every_branch_is_exercised():
    count = collections.Counter()
    for t, r, k, seed in [(48, 24, 12, 11), (12, 3, 6, 12), (30, 4, 10, 13)]:
        count += run_model(t, r, k, seed, nth=3, nj=24)
    # records with three or more non-pivot columns take the fallback path
    count += run_model(12, 6, 6, 14, nth=2, nj=24, only_fallback=True)
    for kind in ("A", "B", "B0", "B1", "C", "vanish", "restrict", "restrict2", "batch", "fallback"):
        assert count[kind] > 0, (kind, dict(count))


@pytest.mark.parametrize("t,r,k,seed", [(24, 12, 6, 21), (40, 10, 5, 22), (60, 20, 4, 23), (9, 2, 4, 24),
                                        (48, 1, 6, 25)])
def test_exact_mode_records_general_restrictions(t, r, k, seed):
    count = run_exact_model(t, r, k, seed, ni=10, nj=12)
    assert count["records"] > 0 and count["rg"] > 0


def test_general_restriction_branches():
    count = collections.Counter()
    for t, r, k, seed in [(24, 12, 6, 31), (40, 10, 5, 32), (16, 4, 8, 33), (30, 15, 6, 34)]:
        count += run_exact_model(t, r, k, seed, ni=12, nj=16)
    # a restriction whose constraint is empty and unsatisfiable needs the
    # random states' records
    for t, r, k, seed in [(12, 3, 6, 12), (30, 4, 10, 13), (20, 2, 8, 15), (16, 1, 10, 16)]:
        count += run_model(t, r, k, seed, nth=3, nj=24)
    for kind in ("rg", "rg_elim", "rg_c", "rg_jii", "rg_empty", "rg_vanish", "A", "B", "C", "iso_batch",
                 "iso_vars", "iso_vanish", "rg_batch", "rg_batch_cols"):
        assert count[kind] > 0, (kind, dict(count))


def test_exponent_above_64_on_diagonal_terms():
    """A projected |L> state against itself: with t = 48 its record keeps
    up to ~44 free variables that all leave through steps without a
    constraint, so p = n0 + nC reaches ~88 (the exact-mode diagonal terms of
    an uneven batch instance, t=48 r=14 k=7)."""
    from circuitsimulator_amd.workloads import uneven_instances
    _, specs = uneven_instances(140, seed=1003)
    t, r, k, ph, xs, zs, L = specs[139]
    assert (t, r, k) == (48, 14, 7)
    Po = orc.make_projector(ph, xs, zs, t)
    count = collections.Counter()
    pmax = 0
    for i in range(1 << k):
        th = orc.prepL(i, t, L)
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        e, p, m = orc.inner(th, orc.prepL(i, t, L))
        want = (e, p, m % 8) if e else (0, 0, 0)
        xj = xj_of(i, L, k)
        assert decode(fallback_ip(rec, xj, t, count)) == want, i
        if popc(((1 << t) - 1) & ~rec.act) <= 2:
            assert decode(unpack(batch_ip(batch_consts(rec, t), xj, count))) == want, i
        assert decode(unpack(batch_ip_gen(batch_consts(rec, t), rec, xj, count))) == want, i
        if e:
            pmax = max(pmax, rec.k + popc(xj) + p)  # n0 + nC of the sweep
    # the case that needs the exponent field's base above 64
    assert pmax > 64


def _seeded_exact_instance(seed):
    """The instance a seed draws: t ~ U{10..39}, r ~ U{1..t/2}, k ~ U{3..8},
    then the projector and L."""
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(seed)
    t = int(rng.integers(10, 40))
    r = int(rng.integers(1, t // 2 + 1))
    k = int(rng.integers(3, 9))
    ph, xs, zs = commuting_paulis(rng, t, r)
    return t, r, k, ph, xs, zs, random_L(rng, t, k)


def test_empty_first_restriction_batch_can_vanish():
    """Record 7 against phi_0 of the seed-200 instance: its first restriction
    has an empty constraint, and another column that misses the free
    variables has right side 1, so the batch ends the term at zero."""
    t, r, k, ph, xs, zs, L = _seeded_exact_instance(200)
    assert (t, r, k) == (11, 4, 4)
    th = orc.prepL(7, t, L)
    assert orc.project(th, orc.make_projector(ph, xs, zs, t)) >= 0
    rec = to_record(th)
    count = collections.Counter()
    got = decode(unpack(batch_ip_gen(batch_consts(rec, t), rec, xj_of(0, L, k), count)))
    e, p, m = orc.inner(th, orc.prepL(0, t, L))
    assert got == ((e, p, m % 8) if e else (0, 0, 0)) == (0, 0, 0)
    assert count["rg_batch_vanish"] == 1


def project_rref(t, xv, gens, count):
    """The record of Pi prepL(x) built directly in reduced row echelon form
    (rows at their pivot lanes), as k_records does for exact-mode records:
    the product state has act = supp(x) and G_a = e_a; each generator
    i^m Z(zeta) X(xi) (the cases of measurePauli, stabilizer.c:827-959) then
    works on the coordinates c = xi & act of xi in that basis.  Returns (the
    record as a Form, sexp), or (None, -1) when the projection vanishes."""
    s = Form()
    s.act = xv & ((1 << t) - 1)
    for a in bits(s.act):
        s.G[a] = 1 << a
    sexp = 0
    for m, zeta, xi in gens:
        act = s.act
        c = xi & act
        rho = xi
        for a in bits(c):
            rho ^= s.G[a]
        vze = sum(1 << a for a in bits(act) if popc(s.G[a] & zeta) & 1)
        w = 2 * m + 4 * (popc(zeta & s.h) & 1) + sum(s.d[a] for a in bits(c))
        w += 4 * sum(1 for a in bits(c) for b in bits(c) if a < b and (s.J[a] >> b) & 1)
        w %= 8
        eta = 0
        for b in bits(act):
            jb = (s.J[b] & ~(1 << b)) | (((s.d[b] >> 1) & 1) << b)
            if (popc(jb & c) + ((vze >> b) & 1)) & 1:
                eta |= 1 << b
        if rho == 0 and w in (0, 4):
            beta = w >> 2
            if eta == 0:
                count["pj_same" if not beta else "pj_vanish"] += 1
                if beta:
                    return None, -1
                continue
            count["pj_shrink"] += 1
            i = low(eta)
            M, di = basis_change(s, i, eta & ~(1 << i), True)
            if beta:
                s.Q = (s.Q + di) % 8
                for x in bits(M):
                    s.d[x] ^= 4
                s.h ^= s.G[i]
            s.act &= ~(1 << i)
        elif rho == 0:
            count["pj_phase"] += 1
            sigma = 1 if w == 2 else -1
            s.Q = (s.Q + sigma) % 8
            for b in bits(eta):
                s.d[b] = (s.d[b] - 2 * sigma) % 8
                s.J[b] ^= eta
        else:
            count["pj_extend"] += 1
            p = low(rho)
            Dp = (w + 4 * (popc(zeta & rho) & 1)) % 8
            for b in bits(act):
                s.J[b] = (s.J[b] & ~(1 << p)) | (((eta >> b) & 1) << p)
            s.G[p], s.d[p] = rho, Dp
            s.J[p] = eta | (((Dp >> 1) & 1) << p)
            s.act |= 1 << p
            T = sum(1 << b for b in bits(s.act) if b != p and (s.G[b] >> p) & 1)
            if T:
                basis_change(s, p, T, True)
        sexp += 1
    r = Form()
    r.act, r.h, r.Q, r.k = s.act, s.h, s.Q % 8, popc(s.act)
    for a in bits(s.act):
        r.G[a] = s.G[a]
        r.d[a] = s.d[a] % 8
        r.J[a] = s.J[a] & s.act
    return r, sexp


def _gens_of(ph, xs, zs):
    return [(int(p) & 3, int(z), int(x)) for p, x, z in zip(ph, xs, zs)]


@pytest.mark.parametrize("t,r,k,seed", [(24, 12, 6, 41), (40, 18, 5, 42), (60, 20, 4, 43), (9, 4, 4, 44),
                                        (48, 1, 6, 45), (12, 6, 3, 46)])
def test_direct_rref_projection_matches_oracle(t, r, k, seed):
    # the directly built record: same projection exponent as the oracle's
    # measurePauli sequence, and every sampled term through the exact-mode
    # record path equal to the oracle's inner product
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    Po = orc.make_projector(ph, xs, zs, t)
    gens = _gens_of(ph, xs, zs)
    count = collections.Counter()
    chi = 1 << k
    js = sorted(set(int(v) for v in rng.integers(0, chi, size=10)))
    for i in range(chi):
        th = orc.prepL(i, t, L)
        e = orc.project(th, Po)
        rec, se = project_rref(t, xj_of(i, L, k), gens, count)
        assert se == e, (i, se, e)
        if e < 0:
            continue
        C = batch_consts(rec, t)
        for j in js:
            ee, p, m = orc.inner(th, orc.prepL(j, t, L))
            want = (ee, p, m % 8) if ee else (0, 0, 0)
            got = decode(unpack(batch_ip_gen(C, rec, xj_of(j, L, k), count)))
            assert got == want, (i, j, got, want)
    assert count["pj_extend"] + count["pj_phase"] + count["pj_shrink"] > 0


def test_direct_rref_projection_branches():
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    count = collections.Counter()
    for seed in range(50, 62):
        rng = np.random.default_rng(seed)
        t = int(rng.integers(6, 30))
        r = int(rng.integers(1, t // 2 + 1))
        k = int(rng.integers(2, 6))
        ph, xs, zs = commuting_paulis(rng, t, r)
        L = random_L(rng, t, k)
        Po = orc.make_projector(ph, xs, zs, t)
        for i in range(1 << k):
            rec, se = project_rref(t, xj_of(i, L, k), _gens_of(ph, xs, zs), count)
            assert se == orc.project(orc.prepL(i, t, L), Po)
    # Z-type generators (all commute; phase 0 or 2) on qubits outside supp(x)
    # leave the product state unchanged or annihilate it
    for seed in range(62, 70):
        rng = np.random.default_rng(seed)
        t, k = 12, 3
        L = [int(v) for v in rng.integers(1, 1 << t, size=k)]
        zs = [int(v) for v in rng.integers(1, 1 << t, size=4)] + [1 << int(q) for q in rng.integers(0, t, size=3)]
        ph = [2 * int(b) for b in rng.integers(0, 2, size=len(zs))]
        xs = [0] * len(zs)
        Po = orc.make_projector(ph, xs, zs, t)
        for i in range(1 << k):
            th = orc.prepL(i, t, L)
            rec, se = project_rref(t, xj_of(i, L, k), _gens_of(ph, xs, zs), count)
            e = orc.project(th, Po)
            assert se == e, (seed, i)
            if e < 0:
                continue
            C = batch_consts(rec, t)
            for j in range(1 << k):
                ee, pp, mm = orc.inner(th, orc.prepL(j, t, L))
                want = (ee, pp, mm % 8) if ee else (0, 0, 0)
                assert decode(unpack(batch_ip_gen(C, rec, xj_of(j, L, k), count))) == want, (seed, i, j)
    for kind in ("pj_extend", "pj_phase", "pj_shrink", "pj_same", "pj_vanish"):
        assert count[kind] > 0, (kind, dict(count))


def h_ops(J, G, h, act, W, cbits, pbits, t, count):
    """phiH_batch's pair operations on a working state, in place on the row
    lists J and G: constraints for the set bits of cbits, then couplings
    for the bits of pbits, each from the lowest bit (pair q = size-1-bit,
    stateprep.c:36-81: bit 1 makes it |00> + |11>, the constraint
    x_2q = x_2q+1, bit 0 the four-term state, the coupling 4 x_2q x_2q+1;
    for odd t the last qubit is |0> on bit 1 and |+> on bit 0).  A
    constraint is a restriction on the working G rows and h (as in
    batch_ip_gen); a coupling, on the working rows, is Q += 4 h_u h_v,
    D += 4 on (cu & hv) ^ (cv & hu) ^ (cu & cv), J ^= cv on cu and cu on cv
    (cu, cv the column ballots of u and v).  Returns (h, act, W), or None
    when a constraint is empty with right side 1 (the term vanishes)."""
    size = (t + 1) // 2
    odd = t & 1
    for b in bits(cbits):
        q = size - 1 - b
        xi = (1 << (t - 1)) if (odd and q == size - 1) else (3 << (2 * q))
        count["h_restrict"] += 1
        col = sum(1 << x for x in range(64) if popc(G[x] & xi) & 1)
        S = col & act
        c = popc(h & xi) & 1
        if not S:
            if c:
                count["h_vanish"] += 1
                return None
            continue
        i = low(S)
        act &= ~(1 << i)
        Gi = G[i]
        for x in bits(S):
            G[x] ^= Gi
        if c:
            h ^= Gi
        W = restrict_step(J, S, c, W)
    for b in bits(pbits):
        q = size - 1 - b
        u, v = 2 * q, 2 * q + 1
        count["h_couple"] += 1
        cu = sum(1 << x for x in range(64) if (G[x] >> u) & 1) & act
        cv = sum(1 << x for x in range(64) if (G[x] >> v) & 1) & act
        hu, hv = (h >> u) & 1, (h >> v) & 1
        if hu & hv:
            W = (W + 0x80000000) & U32
        lin = (cv if hu else 0) ^ (cu if hv else 0) ^ (cu & cv)
        for x in bits(lin):
            J[x] ^= 1 << 63
        for x in bits(cu):
            J[x] ^= cv
        for x in bits(cv):
            J[x] ^= cu
    return h, act, W


def h_masks(t):
    size = (t + 1) // 2
    smask = (1 << size) - 1
    return smask, smask & ~(1 if t & 1 else 0)


def prefix_H(r, jb, vary, t, count):
    """phiH_prefix: the pair operations every term of a batch shares, those
    of the bits of jb outside vary, on the record's rows; W starts at
    128 (k + t - n0) + 1 with Q in m, less 128 per shared constraint.
    Returns the working state (J, G, h, act, W), or None when a shared
    constraint vanishes (every term of the batch is zero)."""
    smask, pairs = h_masks(t)
    J, G, h = packed_rows(r), r.G[:], r.h
    act = r.act | (1 << END) | (1 << TOP)
    cbits = jb & smask & ~vary
    W = (((r.Q & 7) << 29) + 128 * (r.k + t - popc(r.act)) + 1 - 128 * popc(cbits)) & U32
    st = h_ops(J, G, h, act, W, cbits, pairs & ~jb & ~vary, t, count)
    if st is None:
        return None
    h, act, W = st
    return J, G, h, act, W


def batch_ip_H_from(state, j, vary, t, count, iso=True):
    """One term of phiH_batch from a batch's shared state: the pair
    operations of the bits of j in vary, then the loop; returns the parked
    W word."""
    if state is None:
        return 0
    smask, pairs = h_masks(t)
    J0, G0, h, act, W = state
    J, G = J0[:], G0[:]
    cbits = j & smask & vary
    W = (W - 128 * popc(cbits)) & U32
    st = h_ops(J, G, h, act, W, cbits, pairs & ~j & vary, t, count)
    if st is None:
        return 0
    h, act, W = st
    s56, W = elim_loop(J, act, W, count, iso=True if iso else False)
    return W if s56 == END else 0


def extend_H(state, jg, bits, t, count):
    """A group's shared pair operations on top of its batch's state: those of
    the bits of jg in `bits` (the group's shared bits that the batch's terms
    do not share), as phiH_prefix applies them; None when the batch's state
    already vanished or a group constraint vanishes."""
    if state is None:
        return None
    smask, pairs = h_masks(t)
    J0, G0, h, act, W = state
    J, G = J0[:], G0[:]
    cbits = jg & smask & bits
    W = (W - 128 * popc(cbits)) & U32
    st = h_ops(J, G, h, act, W, cbits, pairs & ~jg & bits, t, count)
    if st is None:
        return None
    h, act, W = st
    return J, G, h, act, W


def batch_ip_H_bits(state, j, t, count, iso=True):
    """One term of a group of 8 (phiH_group: bits 0..2 of j its own), the
    three bits in ascending order, each a constraint (set) or a coupling
    (clear, a pair bit), then the loop; returns the parked W word."""
    if state is None:
        return 0
    smask, pairs = h_masks(t)
    J0, G0, h, act, W = state
    J, G = J0[:], G0[:]
    W = (W - 128 * popc(j & smask & 7)) & U32
    for b in range(3):
        bit_b = 1 << b
        st = h_ops(J, G, h, act, W, j & smask & bit_b, pairs & ~j & bit_b, t, count)
        if st is None:
            return 0
        h, act, W = st
    s56, W = elim_loop(J, act, W, count, iso=True if iso else False)
    return W if s56 == END else 0


def batch_ip_H(r, j, t, count, iso=True):
    """One |H^t> term of phiH_batch with every pair operation its own: W
    starts at 128 (k + t - #constraints - n0) + 1 with Q in m.  Returns the
    parked W word."""
    smask, _ = h_masks(t)
    return batch_ip_H_from(prefix_H(r, 0, smask, t, count), j, smask, t, count, iso=iso)


def batch_vary(jb, nb):
    """the bits of j that vary over [jb, jb + nb): everything below the
    highest bit where jb and jb + nb - 1 differ"""
    d = jb ^ (jb + nb - 1)
    return (2 << (d.bit_length() - 1)) - 1 if d else 0


@pytest.mark.parametrize("t,r,seed", [(13, 6, 81), (14, 5, 82), (16, 7, 83), (17, 4, 84), (20, 9, 85),
                                      (21, 8, 86)])
def test_phiH_batch_prefix_matches_oracle(t, r, seed):
    # the shared pair operations of a batch (the bits of j its terms agree
    # on) applied once, then each term's own: every term equals the oracle,
    # for aligned batches of 64, an unaligned one (an exact-mode diagonal
    # start), a short one and a single term
    from circuitsimulator_amd.workloads import commuting_paulis
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r)
    Po = orc.make_projector(ph, xs, zs, t)
    chi = 1 << ((t + 1) // 2)
    thr = orc.thresholds(t)
    count = collections.Counter()
    states = [orc.prepH(int(i), t) for i in rng.integers(0, chi, size=3)]
    states += [orc.random_state(t, seed, s, thr) for s in range(3)]
    shared = 0
    for th in states:
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        for jb, nb in [(64 * int(rng.integers(0, chi // 64)), 64), (int(rng.integers(0, chi - 64)), 64),
                       (64 * int(rng.integers(0, chi // 64)) + 5, 37), (int(rng.integers(0, chi)), 1)]:
            vary = batch_vary(jb, nb)
            shared += popc(((1 << ((t + 1) // 2)) - 1) & ~vary)
            st = prefix_H(rec, jb, vary, t, count)
            for j in range(jb, jb + nb):
                e, p, m = orc.inner(th, orc.prepH(j, t))
                want = (e, p, m % 8) if e else (0, 0, 0)
                assert decode(unpack(batch_ip_H_from(st, j, vary, t, count))) == want, (t, jb, nb, j)
    assert shared > 0 and count["h_restrict"] > 0 and count["h_couple"] > 0


@pytest.mark.parametrize("t,r,seed", [(13, 6, 91), (16, 7, 92), (19, 5, 93), (22, 9, 94)])
def test_phiH_two_level_prefix_matches_oracle(t, r, seed):
    # an aligned batch of 64: the bits above the six low ones once per
    # batch, bits 3..5 once per group of 8, bits 0..2 per term
    from circuitsimulator_amd.workloads import commuting_paulis
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r)
    Po = orc.make_projector(ph, xs, zs, t)
    chi = 1 << ((t + 1) // 2)
    thr = orc.thresholds(t)
    count = collections.Counter()
    states = [orc.prepH(int(i), t) for i in rng.integers(0, chi, size=2)]
    states += [orc.random_state(t, seed, s, thr) for s in range(2)]
    for th in states:
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        jb = 64 * int(rng.integers(0, chi // 64))
        st = prefix_H(rec, jb, 63, t, count)
        for g in range(8):
            sg = extend_H(st, jb + 8 * g, 0x38, t, count)
            for j in range(jb + 8 * g, jb + 8 * g + 8):
                e, p, m = orc.inner(th, orc.prepH(j, t))
                want = (e, p, m % 8) if e else (0, 0, 0)
                assert decode(unpack(batch_ip_H_from(sg, j, 7, t, count))) == want, (t, jb, g, j)
                # the same term with its three pair operations in bit order
                assert decode(unpack(batch_ip_H_bits(sg, j, t, count))) == want, ("bits", t, jb, g, j)
    assert count["h_restrict"] > 0 and count["h_couple"] > 0


def test_phiH_batch_prefix_can_vanish():
    # a projector holding -Z_0 Z_1 fixes x_0 + x_1 = 1, so every j whose
    # highest pair bit (pair 0) is set gives zero: in an aligned batch that
    # bit is shared, and the batch's prefix vanishes for all its terms
    t = 16
    chi = 1 << ((t + 1) // 2)
    ph, xs, zs = [2, 0, 0], [0, 0, 0b110000], [0b11, 0b1100, 0]
    Po = orc.make_projector(ph, xs, zs, t)
    count = collections.Counter()
    vanished = 0
    for i in (3, 77, 200):
        th = orc.prepH(i, t)
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        for jb in range(0, chi, 64):
            st = prefix_H(rec, jb, 63, t, count)
            vanished += st is None
            for j in range(jb, jb + 64, 7):
                e, p, m = orc.inner(th, orc.prepH(j, t))
                want = (e, p, m % 8) if e else (0, 0, 0)
                assert decode(unpack(batch_ip_H_from(st, j, 63, t, count))) == want, (i, j)
    assert vanished > 0


def test_phiH_group_prefix_can_vanish():
    # -Z_8 Z_9 fixes x_8 + x_9 = 1: at t = 16 pair 4 is bit 3 of j, shared by
    # a group of 8 but not by the whole batch, so groups with bit 3 set
    # vanish in their group prefix
    t = 16
    chi = 1 << ((t + 1) // 2)
    Po = orc.make_projector([2, 0], [0, 0b11], [0b11 << 8, 0], t)
    count = collections.Counter()
    vanished = 0
    for i in (5, 90, 201):
        th = orc.prepH(i, t)
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        for jb in range(0, chi, 64):
            st = prefix_H(rec, jb, 63, t, count)
            for g in range(8):
                sg = extend_H(st, jb + 8 * g, 0x38, t, count)
                vanished += sg is None
                for j in range(jb + 8 * g, jb + 8 * g + 8, 3):
                    e, p, m = orc.inner(th, orc.prepH(j, t))
                    want = (e, p, m % 8) if e else (0, 0, 0)
                    assert decode(unpack(batch_ip_H_from(sg, j, 7, t, count))) == want, (i, j)
    assert vanished > 0


@pytest.mark.parametrize("t,r,seed", [(6, 3, 71), (7, 2, 72), (12, 5, 73), (13, 6, 74), (20, 6, 75), (9, 4, 76)])
def test_phiH_terms_match_oracle(t, r, seed):
    # exact-mode records (Pi prepH(i)) and random-state records against every
    # prepH(j): the |H^t> batch form equals the oracle's inner product
    from circuitsimulator_amd.workloads import commuting_paulis
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r)
    Po = orc.make_projector(ph, xs, zs, t)
    chi = 1 << ((t + 1) // 2)
    js = list(range(chi)) if chi <= 64 else [int(v) for v in rng.integers(0, chi, size=40)]
    phis = {j: orc.prepH(j, t) for j in js}
    count = collections.Counter()
    thr = orc.thresholds(t)
    states = [orc.prepH(int(i), t) for i in rng.integers(0, chi, size=6)]
    states += [orc.random_state(t, seed, s, thr) for s in range(6)]
    for th in states:
        if orc.project(th, Po) < 0:
            continue
        rec = to_record(th)
        for j in js:
            e, p, m = orc.inner(th, phis[j])
            want = (e, p, m % 8) if e else (0, 0, 0)
            for iso in (True, False):
                got = decode(unpack(batch_ip_H(rec, j, t, count, iso=iso)))
                assert got == want, (t, j, iso, got, want)
    assert count["h_restrict"] > 0 and count["h_couple"] > 0


def test_phiH_branches():
    from circuitsimulator_amd.workloads import commuting_paulis
    count = collections.Counter()
    for t, r, seed in [(6, 3, 71), (7, 2, 72), (12, 5, 73), (13, 6, 74), (20, 6, 75), (9, 4, 76)]:
        rng = np.random.default_rng(seed)
        ph, xs, zs = commuting_paulis(rng, t, r)
        Po = orc.make_projector(ph, xs, zs, t)
        thr = orc.thresholds(t)
        for s in range(6):
            th = orc.random_state(t, seed, s, thr)
            if orc.project(th, Po) < 0:
                continue
            rec = to_record(th)
            for j in range(min(64, 1 << ((t + 1) // 2))):
                e, p, m = orc.inner(th, orc.prepH(j, t))
                assert decode(unpack(batch_ip_H(rec, j, t, count))) == ((e, p, m % 8) if e else (0, 0, 0))
    for kind in ("h_restrict", "h_couple", "h_vanish", "A", "B", "C", "iso_batch"):
        assert count[kind] > 0, (kind, dict(count))


def wave_find_instance(off, n, key):
    """find_instance (cse_kernels.hip): lane x of each round reads
    off[lo + x*step]; the ballot's highest set lane (lane 0 always set)
    narrows [lo, lo + len) to one step.  Returns (index, rounds)."""
    lo, length, rounds = 0, n, 0
    while length > 1:
        step = (length + 63) >> 6
        ball = 1
        for x in range(64):
            if x * step < length and off[lo + x * step] <= key:
                ball |= 1 << x
        hl = ball.bit_length() - 1
        lo += hl * step
        length = min(step, length - hl * step)
        rounds += 1
    return lo, rounds


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 100, 2048, 4096, 4097, 70000])
def test_wave_instance_search_matches_bisect(n):
    import bisect
    rng = np.random.default_rng(n)
    off = [0] + [int(v) for v in np.cumsum(rng.integers(1, 50, size=n))]
    keys = [int(v) for v in rng.integers(0, off[-1], size=200)] + [0, off[-1] - 1] + off[:min(n, 80)]
    worst = 0
    for key in keys:
        got, rounds = wave_find_instance(off, n, key)
        assert got == bisect.bisect_right(off, key) - 1, (n, key)
        worst = max(worst, rounds)
    # two dependent loads up to 4096 instances, three up to 262144
    assert worst <= (0 if n == 1 else 1 if n <= 64 else 2 if n <= 4096 else 3)



def algorithmic_work(t, r, k, seed, nth, nj):
    """Mean word operations per inner product of the fixation/elimination
    algorithm (the cost model above) on the sampled-mode workload of
    sampled_config(t, r, k, seed): the first nth non-vanishing projected
    random states of the seed's stream, each against nj seeded j, through
    the path the sweep takes (phiL_batch for records with at most two
    non-pivot columns, else the fallback).  Returns (mean, per-kind means,
    per-IP event counts)."""
    from circuitsimulator_amd.workloads import sampled_config
    _, (ph, xs, zs), L = sampled_config(t, r, k, seed=seed)
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    rng = np.random.default_rng(seed + 1)
    js = [int(v) for v in rng.integers(0, 1 << k, size=nj)]
    count = collections.Counter()
    ips, st, done = 0, 0, 0
    while done < nth:
        th = orc.random_state(t, seed, st, thr)
        st += 1
        if orc.project(th, Po) < 0:
            continue
        done += 1
        rec = to_record(th)
        C = batch_consts(rec, t)
        batch = popc(((1 << t) - 1) & ~rec.act) <= 2
        for j in js:
            xj = xj_of(j, L, k)
            if batch:
                batch_ip(C, xj, count)
            else:
                fallback_ip(rec, xj, t, count)
            ips += 1
    kinds = {key: v / ips for key, v in count.items() if key.startswith("w_")}
    events = {key: v / ips for key, v in count.items() if not key.startswith("w_")}
    return sum(kinds.values()), kinds, events


def test_algorithmic_work_of_the_headline_workload():
    # the per-IP figure the benchmark's roofline uses: the t=48, r=24, k=12
    # norm estimation (seed 1004), 64 projected random states x 512 seeded j
    import bench
    mean, kinds, events = algorithmic_work(48, 24, 12, 1004, 64, 512)
    assert round(mean, 1) == bench.ALG_WORDOPS_PER_IP
    # the step statistics behind it: about 7.6 A and 7.1 B steps per IP
    assert 7.0 < events["A"] < 8.3 and 6.5 < events["B"] < 7.8
    assert kinds["w_fix"] == pytest.approx(141.5, abs=0.5)
