"""CPU model of the sweep's record path, checked against the oracle.

The sweep kernel (circuitsimulator_amd/csrc/cse_kernels.hip, cse_device.h)
computes <phi_j|theta'> from an elimination-ready record of theta': the basis
of K in reduced row echelon form with row a stored at its pivot column, the
quadratic form on it, and D packed into bits 61..63 of the J rows.  For
phi_j = prepL(j) the pivot coordinates outside supp(x_j) are fixed in one
batch, the other zero coordinates are restrictions, and the remaining
Gaussian sum is eliminated one variable at a time by a hand-scheduled loop.

This file restates that path in Python, step for step and, for the loop,
register for register (the W word: m * 2^29 | vanished << 28 | nC, and the
accumulator of D_a << 29 plus 2^32 per constrained step), and checks every
term against the oracle's innerProduct (libcirc/stabilizer/stabilizer.c:
589-667).  It also counts the step kinds so that every branch of the loop is
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


def expsum_loop(s, count):
    """The hand-scheduled elimination loop on packed rows J_x | D_x << 61."""
    J = [(s.J[x] & MASK61) | ((s.d[x] & 7) << 61) for x in range(64)]
    act = s.act
    if not act:
        return 1, 0, s.Q & 7
    n0 = popc(act)
    W = ((s.Q + 2 * n0) << 29) & U32
    acc = 0
    while act:
        a = low(act)
        act &= ~(1 << a)
        Ja = J[a]
        hi = Ja >> 32
        if (hi >> 30) & 1:                       # D_a in {2,6}
            count["A"] += 1
            eta = Ja & act
            dm = hi & 0xE0000000
            for x in bits(eta):
                lo = (J[x] & U32) ^ (eta & U32)
                h2 = (((J[x] >> 32) ^ (eta >> 32)) - dm) & U32
                J[x] = (h2 << 32) | lo
            acc += dm
            continue
        eta = Ja & act
        if not eta:                              # no constraint
            count["C"] += 1
            W = (W + 0xC0000001) & U32
            if (hi >> 31) & 1:
                count["vanish"] += 1
                W |= 0x10000000
                break
            continue
        count["B"] += 1                          # constraint eta . z = D_a/4
        i = low(eta)
        act &= ~(1 << i)
        Jr = J[i]
        dmi = (Jr >> 32) & 0xE0000000
        Jc = Jr & MASK61
        for x in range(64):
            lo, h2 = J[x] & U32, J[x] >> 32
            if (eta >> x) & 1:
                lo ^= Jc & U32
                h2 ^= Jc >> 32
                h2 = (h2 + dmi + (0x80000000 if (Jc >> x) & 1 else 0)) & U32
            J[x] = (h2 << 32) | lo
        Mm = Jc ^ (eta if (dmi >> 30) & 1 else 0)
        sp = eta | ((1 << 63) if (hi >> 31) & 1 else 0)
        for x in bits(Mm):
            J[x] ^= sp
        acc += 1 << 32
        if (hi >> 31) & 1:
            W = (W + dmi) & U32
    if W & 0x10000000:
        return 0, 0, 0
    return 1, n0 + (W & 0x0FFFFFFF), ((W >> 29) - (acc >> 30)) & 7


def ip_phiL(r, xj, n, count):
    """Fixation of the pivot coordinates outside supp(x_j), restrictions on
    the other zero coordinates, then the loop; returns (eps, p, m)."""
    F = r.act & ~xj
    B = F & r.h
    s = r.copy()
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
        beta = ((s.h >> q) ^ popc(col & B)) & 1
        if not S:
            if beta:
                return 0, 0, 0
            continue
        i = low(S)
        M, di = basis_change(s, i, S & ~(1 << i), True)
        if beta:
            s.Q = (s.Q + di) & 7
            for x in bits(M):
                s.d[x] ^= 4
            s.h ^= s.G[i]
        s.act &= ~(1 << i)
    eps, pp, mm = expsum_loop(s, count)
    if not eps:
        return 0, 0, 0
    return 1, pp - r.k - popc(xj), mm


def xj_of(j, L, k):
    v = 0
    for rr in range(k):
        if (j >> (k - 1 - rr)) & 1:
            v ^= L[rr]
    return v


def run_model(t, r_gens, k, seed, nth, nj):
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
        for j in range(nj):
            got = ip_phiL(rec, xj_of(j, L, k), t, count)
            e, p, m = orc.inner(th, phis[j])
            want = (e, p, m % 8) if e else (0, 0, 0)
            assert got == want, (t, st - 1, j, got, want)
            checked += 1
    return count


@pytest.mark.parametrize("t,r,k,seed", [(48, 24, 12, 5), (20, 6, 8, 6), (9, 3, 4, 7), (61, 20, 10, 8)])
def test_record_path_matches_oracle(t, r, k, seed):
    count = run_model(t, r, k, seed, nth=3, nj=min(24, 1 << k))
    # every kind of step of the loop ran at least once over the cases
    assert count["A"] > 0 and count["B"] > 0


def test_every_branch_is_exercised():
    count = collections.Counter()
    for t, r, k, seed in [(48, 24, 12, 11), (12, 3, 6, 12), (30, 4, 10, 13)]:
        count += run_model(t, r, k, seed, nth=3, nj=24)
    for kind in ("A", "B", "C", "vanish", "restrict"):
        assert count[kind] > 0, (kind, dict(count))
