// cse_internal.h -- layouts shared by the kernels and the host side of the
// engine (not part of the public C ABI).
#pragma once
#include <stdint.h>

#ifndef DEV
#define DEV __device__ __forceinline__
#endif

// Projector as kernel argument: generator g is i^phase[g] Z(zs[g]) X(xs[g]).
struct DevProjector {
    int nstabs;
    int phase[64];
    uint64_t xs[64];
    uint64_t zs[64];
};

// Elimination-ready record of a projected state theta', REC_WORDS uint64
// words.  The basis of the active rows is in reduced row echelon form and row
// a is stored at index p_a, its pivot column, so that the active-row mask is
// the set of pivot columns and x_a = h_a + z_a for every active a.
//   w[0] = k | n << 8 | Q << 16 | sexp << 24 | vanished << 40
//   w[1] = h, w[2] = active rows (= pivot columns), w[3] = D1, w[4] = D2
//   w[5 .. 69)     G rows
//   w[69 .. 133)   J rows (symmetric; the diagonal is not used)
#define REC_WORDS 133

// Exact mode enumerates only the items that reach the diagonal: record r has
// the chunks c >= floor(r / ch) of its j (the chunk holding j = r and the
// ones after it), nch - floor(r / ch) items.  Block g (records g ch ..
// g ch + ch - 1) thus has ch (nch - g) items, tri_before(g) items precede
// it, and the last block holds the chi - (nch - 1) ch remaining records with
// one item each.
inline __host__ __device__ long long tri_before(long long ch, long long nch, long long g) {
    return ch * (g * (2 * nch - g + 1) / 2);
}
inline __host__ __device__ long long tri_items(long long chi, long long ch, long long nch) {
    return tri_before(ch, nch, nch - 1) + (chi - (nch - 1) * ch);
}

// Exact-mode work items: slices of CSE_SLICE_J consecutive j of one record.
#ifndef CSE_SLICE_J
#define CSE_SLICE_J 64
#endif

// Exact-mode work queue: counters, one per part of the dynamic range, 128
// bytes (16 words) apart, so a launch's claims spread over the XCDs'
// memory channels instead of one address.
#define CSE_NQUEUES 8
#define CSE_QUEUE_STRIDE 16
#define CSE_QUEUE_WORDS (CSE_NQUEUES * CSE_QUEUE_STRIDE)

// Resident blocks (of 4 waves) per CU of the exact-mode sweep
// k_sweep_packed_all: its launch bound and its persistent grid.
#ifndef CSE_EXACT_BLOCKS_PER_CU
#define CSE_EXACT_BLOCKS_PER_CU 7
// A dynamic claim takes remaining / (waves per part x CSE_GSS_DIV) slices;
// the static first claims cover 1 / CSE_STATIC_DIV of all slices
#ifndef CSE_GSS_DIV
#define CSE_GSS_DIV 2
#endif
// The engine's exact-mode rule takes the eight-wave sweep for launches of at
// least this many 64-j slices per resident wave (32 waves per CU)
#ifndef CSE_ALL8_SLICES_PER_WAVE
#define CSE_ALL8_SLICES_PER_WAVE 16
#endif
#ifndef CSE_STATIC_DIV
#define CSE_STATIC_DIV 2
#endif
// largest flush_terms (and the default): 2^30 terms keep every 32-bit bin
// below 2^31
#define CSE_FLUSH_TERMS_MAX (1ll << 30)
#endif

// Histogram bins per instance: P in [-(64 + 64), 0], 8 phases (the public
// CSE_HIST_BINS).
#define HIST_STRIDE (129 * 8)
// Per-wave LDS histogram: p in [-64, 0], 8 phases.
#define LH_BINS (65 * 8)

struct InstDesc {
    int n;              // state dimension t
    int k;              // rows of L (exact == 0)
    int exact;          // 1: |H^t> decomposition, 0: |L>
    int chunk;          // j per work item (exact-mode batches: per slice, 64)
    long long chi;      // decomposition size
    long long nchunks;  // work items per record
    long long rec_off;  // first record of this instance
    long long term_off; // first per-term output slot
    uint64_t seed;      // mode 0: Philox key of the random states
    long long first;    // mode 0: sample index of the first record
    uint64_t L[64];
};

struct SweepArgs {
    int mode;                 // 0 sampled, 1 exact
    int ninst;
    long long nitems;
    const long long *item_off;  // ninst + 1 prefix of items
    // mode 1: items are slices of up to 64 j (InstDesc.chunk).  Global wave
    // g first takes slices [g first_take, g first_take + first_take) of
    // [0, static_end); the rest are cut into CSE_NQUEUES parts of `part`
    // slices, each handed out from its own counter (queue + q
    // CSE_QUEUE_STRIDE) by guided self-scheduling: a claim takes the
    // part's remaining slices (read from its counter just before) /
    // (waves_per_part x CSE_GSS_DIV), at least 1 and at most take_max
    long long take_max;
    // mode 1: a wave flushes its partial histogram of an instance once it
    // has folded this many terms of it (its LDS bins are 32-bit and a term
    // adds at most 2 to one), so no bin can wrap however long the instance
    long long flush_terms;
    long long first_take;
    long long static_end;
    long long part;
    long long waves_per_part;
    const InstDesc *inst;
    const uint64_t *recs;
    unsigned long long *queue;
    int32_t *terms;           // optional per-term output
    long long *zsum;          // mode 0: 4 per record
    double *kahan;            // mode 0: 2 per record; mode 1: 2 per instance
    unsigned long long *hist; // HIST_STRIDE per instance
    long long hist_stride;
    unsigned long long *ips;  // per instance
    int *err;                 // set to 2 if an inner product left [-64, 0]
};

#ifdef __HIPCC__
namespace cse {

struct Rec {
    static DEV void store_vanished(uint64_t *rec, int n) {
        if (lane_id() == 0) {
            rec[0] = ((uint64_t)n << 8) | (1ull << 40);
            rec[1] = 0;
            rec[2] = 0;
            rec[3] = 0;
            rec[4] = 0;
        }
    }
};

struct RecView {
    u64 G, J;   // per lane
    int d;      // per lane: D_x
    int k, n, Q, sexp, vanished;
    u64 h, act, D1, D2;
    DEV void load(const uint64_t *rec) {
        const int x = lane_id();
        const uint64_t w0 = uniform64(rec[0]);
        k = (int)(w0 & 0xff);
        n = (int)((w0 >> 8) & 0xff);
        Q = (int)((w0 >> 16) & 7);
        sexp = (int)((w0 >> 24) & 0xffff);
        vanished = (int)((w0 >> 40) & 1);
        // the header is wave-uniform: read it into scalar registers
        h = uniform64(rec[1]);
        act = uniform64(rec[2]);
        D1 = uniform64(rec[3]);
        D2 = uniform64(rec[4]);
        if (vanished) return;
        G = rec[5 + x];
        J = rec[69 + x];
        d = 2 * (int)((D1 >> x) & 1ull) + 4 * (int)((D2 >> x) & 1ull);
    }
};

}  // namespace cse
#endif
