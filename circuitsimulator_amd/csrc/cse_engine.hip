// cse_engine.hip -- host side of the engine: context, device buffers,
// launches and the C ABI declared in include/cse.h.  Kernels live in
// cse_kernels.hip (same translation unit).
#include "cse_kernels.hip"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cse.h"
#include "cse_host.h"

static_assert(HIST_STRIDE == CSE_HIST_BINS, "device histogram stride is the public histogram size");

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// pinned host memory: device-to-host copies into it are single DMA
// transfers instead of chunked staging through a driver buffer
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// What a captured batch depends on besides the plan's and the results'
// contents: its sizes, kernels and grids (all functions of these) and every
// buffer address its nodes hold.  A call whose key matches replays the graph.
struct GraphKey {
    int mode, ninst, kernel;
    long long nrec, nitems, take_max;
    size_t plan_bytes, out_bytes;
    const void *ptr[8];
};

struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
};

struct cse_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    int cus = 0;
    char name[128] = {0};
    char err[512] = {0};
    DevBuf s1, s2, out, recs, terms, queue, thr, sexp, L, idx, errflag, lsh;
    // a batch's plan: instance descriptors, item and record offsets and
    // projectors in one allocation, laid out as the pinned staging buffer
    DevBuf plan;
    InstDesc *d_inst = nullptr;
    long long *d_itemoff = nullptr, *d_recoff = nullptr;
    DevProjector *d_projs = nullptr;
    // a batch's results (sums, Kahan pairs, histograms, counts, error flag) in
    // one allocation, laid out as the pinned stage they are copied into
    DevBuf outb;
    size_t out_bytes = 0;
    long long *d_zsum = nullptr;
    double *d_kahan = nullptr;
    unsigned long long *d_hist = nullptr, *d_ips = nullptr;
    int32_t *d_sexp = nullptr;
    int *d_err = nullptr;
    HostBuf hstage;  // results coming back (per-sample sums, histograms)
    HostBuf hup;     // the plan going up
    // the memory the current batched call works in: the buffers above, or
    // the static scratch array (cse_exact_projector_noalloc)
    char *b_plan = nullptr, *b_out = nullptr, *b_hup = nullptr, *b_hstage = nullptr;
    uint64_t *b_recs = nullptr;
    unsigned long long *b_queue = nullptr;
    int32_t *b_sexp1 = nullptr, *b_terms = nullptr;
    // batched calls replayed from captured HIP graphs (cse_set_graphs)
    bool use_graphs = false;
    std::vector<GraphEntry> graphs;
    unsigned long long graph_clock = 0;
    // this device's copy of the static scratch array (nullptr if its address
    // could not be obtained; scratch_err says why)
    char *scratch = nullptr;
    char scratch_err[160] = {0};
    bool thr_ready = false;
    int blocks_per_cu = 8;  // persistent grid: blocks of 4 waves per CU
    int sweep_kernel[2] = {0, 0};  // per mode (cse_set_sweep_kernel)
    // exact mode: terms of one instance after which a wave flushes its
    // partial histogram (CSE_FLUSH_TERMS in the environment, at most and by
    // default CSE_FLUSH_TERMS_MAX)
    long long flush_terms = CSE_FLUSH_TERMS_MAX;
};

// ------------------------------------------------- static scratch array
// cse_exact_projector_noalloc works in this file-scope device array instead
// of allocated buffers: the work queue, the plan, the results block, the
// projection exponents and the records of one instance of chi <= 2^13.  Each
// device holds its own copy for the life of the process; its address is
// taken once per device with hipGetSymbolAddress when the first context on
// that device is created, and host code only ever uses that address.
constexpr int kStaticMaxLogChi = 13;
constexpr long long kStaticMaxChi = 1ll << kStaticMaxLogChi;
constexpr size_t kStaticPlanBytes = 4096;  // one instance's plan: 2,816 B
constexpr size_t kStaticOutBytes = 8448;   // one instance's results block: 8,288 B
constexpr size_t kScratchQueue = 0;
constexpr size_t kScratchPlan = 8 * CSE_QUEUE_WORDS;
constexpr size_t kScratchOut = kScratchPlan + kStaticPlanBytes;
constexpr size_t kScratchSexp = kScratchOut + kStaticOutBytes;
constexpr size_t kScratchRecs = kScratchSexp + 4 * (size_t)kStaticMaxChi;
constexpr size_t kScratchBytes = kScratchRecs + (size_t)REC_WORDS * 8 * (size_t)kStaticMaxChi;  // 8,762,624 B
static_assert(kScratchOut % 256 == 0 && kScratchSexp % 256 == 0 && kScratchRecs % 256 == 0, "scratch parts aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxScratchDevices];
static bool g_scratch_tried[kMaxScratchDevices];
static char g_scratch_msg[kMaxScratchDevices][128];

namespace {

// The message is formatted into a local buffer first: an argument may be
// the context's own previous message (fail(ctx, rc, "...: %s", ctx->err)).
int fail(cse_ctx *c, int code, const char *fmt, ...) {
    char msg[sizeof(c->err)];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    memcpy(c->err, msg, sizeof(msg));
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ctx, CSE_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

int ensure(cse_ctx *ctx, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return CSE_OK;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, CSE_ERR_NOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return CSE_OK;
}

#define ENSURE(buf, bytes)                            \
    do {                                              \
        int rc_ = ensure(ctx, ctx->buf, (bytes));     \
        if (rc_) return rc_;                          \
    } while (0)

int ensure_host(cse_ctx *ctx, HostBuf &b, size_t bytes) {
    if (b.cap >= bytes) return CSE_OK;
    if (b.p) HIPCHK(hipHostFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 4;
    hipError_t e = hipHostMalloc(&b.p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, CSE_ERR_NOMEM, "hipHostMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return CSE_OK;
}

template <typename T>
T *dp(DevBuf &b) { return (T *)b.p; }

int zero64(cse_ctx *ctx, void *p, long long words) {
    if (words <= 0) return CSE_OK;
    int blocks = (int)((words + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero64, dim3(blocks), dim3(256), 0, ctx->stream, (unsigned long long *)p, words);
    HIPCHK(hipGetLastError());
    return CSE_OK;
}

// buffers a batch clears before its launches, in one k_zero_list launch
void zl_add(ZeroList &zl, void *p, long long words) {
    if (words <= 0) return;
    zl.p[zl.count] = (unsigned long long *)p;
    zl.n[zl.count] = words;
    zl.count++;
}

int zero_list(cse_ctx *ctx, const ZeroList &zl) {
    long long most = 0;
    for (int b = 0; b < zl.count; b++) most = std::max(most, zl.n[b]);
    if (most == 0) return CSE_OK;
    const int blocks = (int)std::min<long long>(4096, (most + 255) / 256);
    hipLaunchKernelGGL(k_zero_list, dim3(blocks), dim3(256), 0, ctx->stream, zl);
    HIPCHK(hipGetLastError());
    return CSE_OK;
}

int ensure_thresholds(cse_ctx *ctx) {
    if (ctx->thr_ready) return CSE_OK;
    std::vector<uint64_t> thr(65 * 65, 0);
    for (int n = 1; n <= 64; n++) cse_dimension_thresholds(n, &thr[65 * n]);
    ENSURE(thr, thr.size() * 8);
    HIPCHK(hipMemcpyAsync(ctx->thr.p, thr.data(), thr.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->thr_ready = true;
    return CSE_OK;
}

DevProjector to_dev(const cse_projector *P) {
    DevProjector d;
    memset(&d, 0, sizeof(d));
    if (!P) return d;
    d.nstabs = P->nstabs;
    for (int g = 0; g < P->nstabs && g < 64; g++) {
        d.phase[g] = P->phase[g] & 3;
        d.xs[g] = P->xs[g];
        d.zs[g] = P->zs[g];
    }
    return d;
}

int check_projector(cse_ctx *ctx, const cse_projector *P) {
    if (!P) return fail(ctx, CSE_ERR_ARG, "projector is NULL");
    if (P->nqubits < 1 || P->nqubits > CSE_NMAX)
        return fail(ctx, CSE_ERR_ARG, "projector width t=%d outside [1, %d]", P->nqubits, CSE_NMAX);
    if (P->nstabs < 1 || P->nstabs > CSE_NMAX)
        return fail(ctx, CSE_ERR_ARG, "projector has %d generators (need 1..%d)", P->nstabs, CSE_NMAX);
    uint64_t m = P->nqubits == 64 ? ~0ull : ((1ull << P->nqubits) - 1);
    for (int g = 0; g < P->nstabs; g++)
        if ((P->xs[g] & ~m) || (P->zs[g] & ~m))
            return fail(ctx, CSE_ERR_ARG, "generator %d has bits beyond t=%d", g, P->nqubits);
    return CSE_OK;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

int default_chunk(long long chi, int size = 256) {
    long long c = size;
    if (chi < c) c = chi;
    return (int)(c < 1 ? 1 : c);
}

// Exact mode's default largest claim: 512 j when the launch still has eight
// items of that size per resident wave of k_sweep_packed_all (seven waves per
// SIMD), else 256.  Larger claims halve the per-segment record load, table
// set-up and histogram flush; the guided self-scheduling of the sweep shrinks
// the claims at the end of the queue in any case.
int exact_chunk(cse_ctx *ctx, const cse_instance *in, int ninst) {
    long long items = 0;
    for (int i = 0; i < ninst; i++) {
        const long long chi = in[i].exact ? (1ll << ((in[i].P.nqubits + 1) / 2)) : (1ll << in[i].k);
        // items of record r that reach the diagonal or beyond
        const long long per = (chi + 511) / 512;
        items += chi * (per + 1) / 2;
    }
    const long long waves = (long long)ctx->cus * CSE_EXACT_BLOCKS_PER_CU * 4;
    return items >= 8 * waves ? 512 : 256;
}

// Persistent grid for k_sweep: bpc resident blocks of 4 waves per CU
int sweep_grid(cse_ctx *ctx, long long nitems, int bpc) {
    long long g = (long long)ctx->cus * bpc;
    long long need = (nitems + 3) / 4;
    if (need < g) g = need;
    return (int)(g < 1 ? 1 : g);
}

// Records for instances [0, ninst): mode 0 random states (sample indices
// first + i), mode 1 projected decomposition states.  Then the sweep.
struct Plan {
    int ninst = 0;
    InstDesc *inst = nullptr;
    long long *rec_off = nullptr;   // ninst + 1
    long long *item_off = nullptr;  // ninst + 1
    DevProjector *projs = nullptr;
    long long nrec = 0, nitems = 0, nterms = 0;
    long long take_max = 1;  // exact mode: largest claim, in slices of 64 j
    // appends an instance of nr records, ni work items and nt terms (its
    // record and term offsets are set here)
    void add(InstDesc d, const DevProjector &p, long long nr, long long ni, long long nt) {
        if (ninst == 0) rec_off[0] = item_off[0] = 0;
        d.rec_off = nrec;
        d.term_off = nterms;
        inst[ninst] = d;
        projs[ninst] = p;
        nrec += nr;
        nitems += ni;
        nterms += nt;
        ninst++;
        rec_off[ninst] = nrec;
        item_off[ninst] = nitems;
    }
};

// Heap storage of a plan of up to n instances.
struct PlanStore {
    std::vector<InstDesc> inst;
    std::vector<long long> ro, io;
    std::vector<DevProjector> pj;
    // n = 0 allocates nothing (the static path brings its own storage)
    explicit PlanStore(int n) : inst(n), ro(n ? n + 1 : 0), io(n ? n + 1 : 0), pj(n) {}
    Plan plan() {
        Plan p;
        p.inst = inst.data();
        p.rec_off = ro.data();
        p.item_off = io.data();
        p.projs = pj.data();
        return p;
    }
};

// Plan layout: instance descriptors, item and record offsets, projectors,
// each part 256-byte aligned.  Returns the total size.
size_t plan_layout(const Plan &pl, size_t off[4], size_t sz[4]) {
    const size_t ninst = (size_t)pl.ninst;
    sz[0] = sizeof(InstDesc) * ninst;
    sz[1] = sizeof(long long) * (ninst + 1);
    sz[2] = sizeof(long long) * (ninst + 1);
    sz[3] = sizeof(DevProjector) * ninst;
    size_t total = 0;
    for (int b = 0; b < 4; b++) {
        off[b] = total;
        total += (sz[b] + 255) & ~(size_t)255;
    }
    return total;
}

// The plan is written into the bound host stage in its device layout, and
// the device pointers into the bound plan are set; copy_plan then moves it
// in one transfer.  Every call synchronizes its stream before returning, so
// the stage is free again at the next call.
size_t stage_plan(cse_ctx *ctx, const Plan &pl) {
    size_t off[4], sz[4];
    const size_t total = plan_layout(pl, off, sz);
    const void *src[4] = {pl.inst, pl.item_off, pl.rec_off, pl.projs};
    char *h = ctx->b_hup;
    for (int b = 0; b < 4; b++) memcpy(h + off[b], src[b], sz[b]);
    char *d = ctx->b_plan;
    ctx->d_inst = (InstDesc *)(d + off[0]);
    ctx->d_itemoff = (long long *)(d + off[1]);
    ctx->d_recoff = (long long *)(d + off[2]);
    ctx->d_projs = (DevProjector *)(d + off[3]);
    return total;
}

int copy_plan(cse_ctx *ctx, size_t bytes) {
    HIPCHK(hipMemcpyAsync(ctx->b_plan, ctx->b_hup, bytes, hipMemcpyHostToDevice, ctx->stream));
    return CSE_OK;
}

int run_records(cse_ctx *ctx, const Plan &pl, int mode) {
    const int ninst = pl.ninst;
    const long long waves = pl.nrec;
    const long long blocks = (waves + 3) / 4;
    if (blocks > 0) {
        hipLaunchKernelGGL(k_records, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, mode, ninst,
                           ctx->d_recoff, ctx->d_inst, ctx->d_projs,
                           mode == 0 ? dp<uint64_t>(ctx->thr) : nullptr, pl.nrec, ctx->b_recs, ctx->d_sexp,
                           ctx->d_err);
        HIPCHK(hipGetLastError());
    }
    return CSE_OK;
}

// Layout of a batch's results block, the same on the device and in the
// pinned stage.  Sampled mode (0): per-sample Z[omega] sums (32 B), Kahan
// pairs (16 B) and projection exponents (4 B), then per-instance counts and
// the error flag.  Exact mode (1): per-instance histograms, Kahan pairs and
// counts, then the error flag.
struct OutLayout {
    size_t zsum, kahan, sexp, hist, ips, err, total;
};

OutLayout out_layout(const Plan &pl, int mode) {
    const size_t n = (size_t)pl.nrec, ni = (size_t)pl.ninst;
    OutLayout o;
    memset(&o, 0, sizeof(o));
    if (mode == 0) {
        o.zsum = 0;
        o.kahan = 32 * n;
        o.sexp = 48 * n;
        o.ips = (52 * n + 7) & ~(size_t)7;
        o.err = o.ips + 8 * ni;
    } else {
        const size_t nh = (size_t)HIST_STRIDE * ni;
        o.hist = 0;
        o.kahan = 8 * nh;
        o.ips = o.kahan + 16 * ni;
        o.err = o.ips + 8 * ni;
    }
    o.total = o.err + 8;
    return o;
}

// Binds a batch to the context's own buffers, grown on demand (device memory
// and pinned host stages).
int bind_ctx_mem(cse_ctx *ctx, const Plan &pl, int mode, bool want_terms) {
    size_t off[4], sz[4];
    const size_t pbytes = plan_layout(pl, off, sz);
    const OutLayout o = out_layout(pl, mode);
    ENSURE(plan, pbytes);
    ENSURE(outb, o.total);
    ENSURE(recs, (size_t)REC_WORDS * 8 * (size_t)pl.nrec);
    ENSURE(queue, 8 * CSE_QUEUE_WORDS);
    if (mode == 1) ENSURE(sexp, 4 * (size_t)pl.nrec);
    if (want_terms) ENSURE(terms, 4 * (size_t)(pl.nterms > 0 ? pl.nterms : 1));
    int rc = ensure_host(ctx, ctx->hup, pbytes);
    if (!rc) rc = ensure_host(ctx, ctx->hstage, o.total);
    if (rc) return rc;
    ctx->b_plan = (char *)ctx->plan.p;
    ctx->b_out = (char *)ctx->outb.p;
    ctx->b_recs = dp<uint64_t>(ctx->recs);
    ctx->b_queue = dp<unsigned long long>(ctx->queue);
    ctx->b_sexp1 = mode == 1 ? dp<int32_t>(ctx->sexp) : nullptr;
    ctx->b_terms = want_terms ? dp<int32_t>(ctx->terms) : nullptr;
    ctx->b_hup = (char *)ctx->hup.p;
    ctx->b_hstage = (char *)ctx->hstage.p;
    return CSE_OK;
}

// Binds a one-instance exact-mode batch to this device's static scratch
// array and the caller's host stages; allocates nothing.
int bind_static_mem(cse_ctx *ctx, const Plan &pl, char *hup, char *hstage) {
    if (!ctx->scratch) return fail(ctx, CSE_ERR_HIP, "static scratch array unavailable: %s", ctx->scratch_err);
    size_t off[4], sz[4];
    const size_t pbytes = plan_layout(pl, off, sz);
    const OutLayout o = out_layout(pl, 1);
    if (pl.nrec > kStaticMaxChi)
        return fail(ctx, CSE_ERR_ARG, "chi=%lld above the no-allocation limit 2^%d", pl.nrec, kStaticMaxLogChi);
    if (pbytes > kStaticPlanBytes || o.total > kStaticOutBytes)
        return fail(ctx, CSE_ERR_ARG, "internal: plan or results of %zu / %zu B above the static stages", pbytes,
                    o.total);
    ctx->b_queue = (unsigned long long *)(ctx->scratch + kScratchQueue);
    ctx->b_plan = ctx->scratch + kScratchPlan;
    ctx->b_out = ctx->scratch + kScratchOut;
    ctx->b_sexp1 = (int32_t *)(ctx->scratch + kScratchSexp);
    ctx->b_recs = (uint64_t *)(ctx->scratch + kScratchRecs);
    ctx->b_terms = nullptr;
    ctx->b_hup = hup;
    ctx->b_hstage = hstage;
    return CSE_OK;
}

// One launch that clears the bound results block (every sum, count and flag;
// the exponents are rewritten by the record launch) and the work queue.
int prep_batch(cse_ctx *ctx, const Plan &pl, int mode) {
    const OutLayout o = out_layout(pl, mode);
    char *b = ctx->b_out;
    ctx->out_bytes = o.total;
    ctx->d_zsum = mode == 0 ? (long long *)(b + o.zsum) : nullptr;
    ctx->d_kahan = (double *)(b + o.kahan);
    ctx->d_sexp = mode == 0 ? (int32_t *)(b + o.sexp) : ctx->b_sexp1;
    ctx->d_hist = mode == 1 ? (unsigned long long *)(b + o.hist) : nullptr;
    ctx->d_ips = (unsigned long long *)(b + o.ips);
    ctx->d_err = (int *)(b + o.err);
    ZeroList zl;
    memset(&zl, 0, sizeof(zl));
    zl_add(zl, b, (long long)(o.total / 8));
    zl_add(zl, ctx->b_queue, CSE_QUEUE_WORDS);
    return zero_list(ctx, zl);
}

enum SweepKernel { SK_GENERIC, SK_PACKED, SK_PACKED_GEN, SK_ALL, SK_ALL8 };

// The sweep kernel of a launch.  Exact mode takes k_sweep_packed_all (seven
// waves per SIMD) or k_sweep_packed_all8 (eight) as set, and by the engine's
// rule the eight-wave kernel when the launch holds at least
// CSE_ALL8_SLICES_PER_WAVE slices for each of its waves: a long launch gains
// from the extra wave's latency hiding (the uneven batch 2.5%, the t=30 |H^t>
// projector 2.6%), a short one loses more to its spills than it gains (one
// projector of 2,048 records, 4.6 slices per wave: 4.5%).  Sampled launches take
// k_sweep_packed (eight waves, the two-column asm path for |L> terms), or
// k_sweep_packed_gen when set or when every term is an |H^t> term: that
// kernel runs an aligned batch of |H^t> terms as eight groups of 8 that share
// the pair operations of bits 3..5 of j.  Records of n > 60 take k_sweep.
SweepKernel sweep_choice(const cse_ctx *ctx, const Plan &pl, int mode) {
    bool packed = true, allH = true;
    for (int i = 0; i < pl.ninst; i++) {
        packed = packed && pl.inst[i].n <= CSE_END_ROW;
        allH = allH && pl.inst[i].exact;
    }
    if (!packed) return SK_GENERIC;
    if (mode == 1) {
        const int k = ctx->sweep_kernel[1];
        if (k) return k == 2 ? SK_ALL8 : SK_ALL;
        return pl.nitems >= CSE_ALL8_SLICES_PER_WAVE * 32ll * ctx->cus ? SK_ALL8 : SK_ALL;
    }
    const int k = ctx->sweep_kernel[0];
    return k == 2 || (k == 0 && allH) ? SK_PACKED_GEN : SK_PACKED;
}

int run_sweep(cse_ctx *ctx, const Plan &pl, int mode, bool want_terms) {
    const int ninst = pl.ninst;
    SweepArgs A;
    A.mode = mode;
    A.ninst = ninst;
    A.nitems = pl.nitems;
    A.item_off = ctx->d_itemoff;
    A.inst = ctx->d_inst;
    A.recs = ctx->b_recs;
    A.queue = ctx->b_queue;
    A.terms = want_terms ? ctx->b_terms : nullptr;
    A.zsum = ctx->d_zsum;
    A.kahan = ctx->d_kahan;
    A.hist = ctx->d_hist;
    A.hist_stride = HIST_STRIDE;
    A.ips = ctx->d_ips;
    A.err = ctx->d_err;
    const SweepKernel sk = sweep_choice(ctx, pl, mode);
    const int grid = sweep_grid(ctx, pl.nitems, sk == SK_ALL8 ? 8 : sk == SK_ALL ? CSE_EXACT_BLOCKS_PER_CU
                                                                                 : ctx->blocks_per_cu);
    A.take_max = pl.take_max;
    A.flush_terms = ctx->flush_terms;
    // the static first claims cover about 1 / CSE_STATIC_DIV of the slices
    // (at most take_max per wave); the rest is cut into CSE_NQUEUES parts
    const long long waves = 4ll * grid;
    A.first_take = pl.nitems / (CSE_STATIC_DIV * waves);
    if (A.first_take > A.take_max) A.first_take = A.take_max;
    if (A.first_take < 1) A.first_take = 1;
    A.static_end = A.first_take * waves < pl.nitems ? A.first_take * waves : pl.nitems;
    A.part = (pl.nitems - A.static_end + CSE_NQUEUES - 1) / CSE_NQUEUES;
    A.waves_per_part = (waves + CSE_NQUEUES - 1) / CSE_NQUEUES;
    switch (sk) {
    case SK_ALL8: hipLaunchKernelGGL(k_sweep_packed_all8, dim3(grid), dim3(256), 0, ctx->stream, A); break;
    case SK_ALL: hipLaunchKernelGGL(k_sweep_packed_all, dim3(grid), dim3(256), 0, ctx->stream, A); break;
    case SK_PACKED_GEN: hipLaunchKernelGGL(k_sweep_packed_gen, dim3(grid), dim3(256), 0, ctx->stream, A); break;
    case SK_PACKED: hipLaunchKernelGGL(k_sweep_packed, dim3(grid), dim3(256), 0, ctx->stream, A); break;
    default: hipLaunchKernelGGL(k_sweep, dim3(grid), dim3(256), 0, ctx->stream, A); break;
    }
    HIPCHK(hipGetLastError());
    return CSE_OK;
}

// The device work of a bound, staged batch in stream order: plan copy,
// clear, records, sweep, results copy; with events between the launches
// when timed.
int enqueue_batch(cse_ctx *ctx, const Plan &pl, int mode, bool want_terms, size_t plan_bytes, bool timed) {
    int rc = copy_plan(ctx, plan_bytes);
    if (!rc) rc = prep_batch(ctx, pl, mode);
    if (rc) return rc;
    if (timed) HIPCHK(hipEventRecord(ctx->e0, ctx->stream));
    rc = run_records(ctx, pl, mode);
    if (rc) return rc;
    if (timed) HIPCHK(hipEventRecord(ctx->e1, ctx->stream));
    rc = run_sweep(ctx, pl, mode, want_terms);
    if (rc) return rc;
    if (timed) HIPCHK(hipEventRecord(ctx->e2, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->b_hstage, ctx->b_out, ctx->out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    return CSE_OK;
}

constexpr size_t kMaxGraphs = 16;

// Runs a bound, staged batch: eagerly, or -- when the context replays graphs
// and the call writes no per-term output -- from the context's graph of the
// same key, captured at the first call of that key.  *graphed tells which;
// the caller synchronizes the stream.
int run_batch(cse_ctx *ctx, const Plan &pl, int mode, bool want_terms, size_t plan_bytes, bool *graphed) {
    *graphed = false;
    if (!ctx->use_graphs || want_terms) return enqueue_batch(ctx, pl, mode, want_terms, plan_bytes, true);
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.mode = mode;
    key.ninst = pl.ninst;
    key.kernel = (int)sweep_choice(ctx, pl, mode);
    key.nrec = pl.nrec;
    key.nitems = pl.nitems;
    key.take_max = pl.take_max;
    key.plan_bytes = plan_bytes;
    key.out_bytes = out_layout(pl, mode).total;
    const void *ptr[8] = {ctx->b_plan, ctx->b_out, ctx->b_recs, ctx->b_queue,
                          ctx->b_sexp1, ctx->b_hup, ctx->b_hstage, mode == 0 ? ctx->thr.p : nullptr};
    memcpy(key.ptr, ptr, sizeof(ptr));
    GraphEntry *hit = nullptr;
    for (GraphEntry &g : ctx->graphs)
        if (memcmp(&g.key, &key, sizeof(key)) == 0) {
            hit = &g;
            break;
        }
    if (!hit) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch(ctx, pl, mode, false, plan_bytes, false);
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return rc ? rc : fail(ctx, CSE_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            return fail(ctx, CSE_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
        if (ctx->graphs.size() >= kMaxGraphs) {
            size_t old = 0;
            for (size_t g = 1; g < ctx->graphs.size(); g++)
                if (ctx->graphs[g].used < ctx->graphs[old].used) old = g;
            (void)hipGraphExecDestroy(ctx->graphs[old].exec);
            ctx->graphs.erase(ctx->graphs.begin() + (long)old);
        }
        GraphEntry ent;
        ent.key = key;
        ent.exec = exec;
        ctx->graphs.push_back(ent);
        hit = &ctx->graphs.back();
    }
    hit->used = ++ctx->graph_clock;
    HIPCHK(hipEventRecord(ctx->e0, ctx->stream));
    HIPCHK(hipGraphLaunch(hit->exec, ctx->stream));
    HIPCHK(hipEventRecord(ctx->e2, ctx->stream));
    *graphed = true;
    return CSE_OK;
}

InstDesc make_inst(int t, int k, int exact, const uint64_t *L, int chunk) {
    InstDesc d;
    memset(&d, 0, sizeof(d));
    d.n = t;
    d.k = exact ? 0 : k;
    d.exact = exact ? 1 : 0;
    d.chi = exact ? (1ll << ((t + 1) / 2)) : (1ll << k);
    d.chunk = default_chunk(d.chi, chunk > 0 ? chunk : 256);
    if (d.chunk > d.chi) d.chunk = (int)d.chi;
    d.nchunks = (d.chi + d.chunk - 1) / d.chunk;
    if (!exact)
        for (int r = 0; r < k; r++) d.L[r] = L[r];
    return d;
}

// chi <= 2^30 keeps every per-theta Z[omega] sum exact in int64: a term is at
// most 2^32 in each component at the 2^32 fixed-point scale, so chi terms stay
// below 2^62.  (The reference holds chi in an int, innerprod.c:122-131.)
constexpr int kMaxLogChi = 30;
// samples (records) per sampled-mode call
constexpr long long kMaxSamples = 1ll << 24;

int check_decomp(cse_ctx *ctx, int t, int k, int exact, const uint64_t *L) {
    if (exact) {
        if ((t + 1) / 2 > kMaxLogChi)
            return fail(ctx, CSE_ERR_ARG, "exact decomposition chi=2^%d exceeds 2^%d", (t + 1) / 2, kMaxLogChi);
        return CSE_OK;
    }
    if (k < 0 || k > kMaxLogChi) return fail(ctx, CSE_ERR_ARG, "k=%d outside [0, %d]", k, kMaxLogChi);
    if (k > 0 && !L) return fail(ctx, CSE_ERR_ARG, "L is NULL");
    uint64_t m = t == 64 ? ~0ull : ((1ull << t) - 1);
    for (int r = 0; r < k; r++)
        if (L[r] & ~m) return fail(ctx, CSE_ERR_ARG, "row %d of L has bits beyond t=%d", r, t);
    return CSE_OK;
}

// the message of the calling thread's last failed cse_create (no context
// exists to hold it); cse_error(NULL) returns it
thread_local char g_create_err[256];

// a state-level call's count (pairs, states) is passed to its kernel as an
// int, and its grid is count / 4 blocks
constexpr size_t kMaxStates = size_t(1) << 30;

int init_ctx(cse_ctx *ctx, int device) {
    ctx->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ctx->e0));
    HIPCHK(hipEventCreate(&ctx->e1));
    HIPCHK(hipEventCreate(&ctx->e2));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    ctx->cus = prop.multiProcessorCount;
    snprintf(ctx->name, sizeof(ctx->name), "%s (%s)", prop.name, prop.gcnArchName);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ctx, CSE_ERR_HIP, "device %d is %s; this engine is built for gfx950", device, prop.gcnArchName);
    if (const char *ft = getenv("CSE_FLUSH_TERMS")) {
        const long long v = atoll(ft);
        if (v < 1 || v > CSE_FLUSH_TERMS_MAX)
            return fail(ctx, CSE_ERR_ARG, "CSE_FLUSH_TERMS=%s outside [1, %lld]", ft, CSE_FLUSH_TERMS_MAX);
        ctx->flush_terms = v;
    }
    // this device's copy of the static scratch array, once per device (with
    // the device current); a failure only disables the no-allocation call
    if (device < kMaxScratchDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_tried[device]) {
            g_scratch_tried[device] = true;
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e == hipSuccess && p) {
                g_scratch_ptr[device] = (char *)p;
            } else {
                (void)hipGetLastError();
                snprintf(g_scratch_msg[device], sizeof(g_scratch_msg[device]), "hipGetSymbolAddress: %s",
                         e != hipSuccess ? hipGetErrorString(e) : "null address");
            }
        }
        ctx->scratch = g_scratch_ptr[device];
        if (!ctx->scratch) snprintf(ctx->scratch_err, sizeof(ctx->scratch_err), "%s", g_scratch_msg[device]);
    } else {
        snprintf(ctx->scratch_err, sizeof(ctx->scratch_err), "device index %d above %d", device,
                 kMaxScratchDevices - 1);
    }
    return CSE_OK;
}

}  // namespace

// ====================================================================== ABI
extern "C" {

int cse_create(int device, cse_ctx **out) {
    g_create_err[0] = 0;
    if (!out) {
        snprintf(g_create_err, sizeof(g_create_err), "cse_create: NULL output pointer");
        return CSE_ERR_ARG;
    }
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        snprintf(g_create_err, sizeof(g_create_err), "cse_create: no HIP device (%s)",
                 e != hipSuccess ? hipGetErrorString(e) : "device count 0");
        return CSE_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        snprintf(g_create_err, sizeof(g_create_err), "cse_create: device %d outside [0, %d)", device, count);
        return CSE_ERR_ARG;
    }
    cse_ctx *ctx = new cse_ctx();
    const int rc = init_ctx(ctx, device);
    if (rc != CSE_OK) {
        snprintf(g_create_err, sizeof(g_create_err), "cse_create: %s", ctx->err);
        cse_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return CSE_OK;
}

void cse_destroy(cse_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (GraphEntry &g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
    DevBuf *bufs[] = {&ctx->s1, &ctx->s2, &ctx->out, &ctx->recs, &ctx->terms, &ctx->queue, &ctx->thr,
                      &ctx->sexp, &ctx->L, &ctx->idx, &ctx->errflag, &ctx->lsh, &ctx->plan, &ctx->outb};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (ctx->hstage.p) (void)hipHostFree(ctx->hstage.p);
    if (ctx->hup.p) (void)hipHostFree(ctx->hup.p);
    if (ctx->e0) (void)hipEventDestroy(ctx->e0);
    if (ctx->e1) (void)hipEventDestroy(ctx->e1);
    if (ctx->e2) (void)hipEventDestroy(ctx->e2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *cse_error(const cse_ctx *ctx) {
    if (ctx) return ctx->err;
    return g_create_err[0] ? g_create_err : "no context";
}

int cse_set_sweep_kernel(cse_ctx *ctx, int mode, int kernel) {
    if (!ctx) return CSE_ERR_ARG;
    if (mode < 0 || mode > 1) return fail(ctx, CSE_ERR_ARG, "sweep mode %d outside [0, 1]", mode);
    if (kernel < 0 || kernel > 2) return fail(ctx, CSE_ERR_ARG, "sweep kernel %d outside [0, 2]", kernel);
    ctx->sweep_kernel[mode] = kernel;
    return CSE_OK;
}

int cse_set_graphs(cse_ctx *ctx, int enable) {
    if (!ctx) return CSE_ERR_ARG;
    ctx->use_graphs = enable != 0;
    return CSE_OK;
}

int cse_device_info(cse_ctx *ctx, int *cus, char *name, int name_len) {
    if (!ctx) return CSE_ERR_ARG;
    if (cus) *cus = ctx->cus;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", ctx->name);
    return CSE_OK;
}

int cse_inner_products(cse_ctx *ctx, int n, size_t npairs, const uint64_t *s1, const uint64_t *s2,
                       int32_t *out, int one_launch_per_pair, double *ms) {
    if (!ctx) return CSE_ERR_ARG;
    if (n < 1 || n > CSE_NMAX) return fail(ctx, CSE_ERR_ARG, "n=%d outside [1, %d]", n, CSE_NMAX);
    if (npairs == 0) {
        if (ms) *ms = 0;
        return CSE_OK;
    }
    if (!s1 || !s2 || !out) return fail(ctx, CSE_ERR_ARG, "NULL buffer");
    if (npairs > kMaxStates) return fail(ctx, CSE_ERR_ARG, "%zu pairs above 2^30 per call", npairs);
    const size_t words = (size_t)CSE_STATE_WORDS(n);
    for (size_t i = 0; i < npairs; i++) {
        const uint64_t *a = s1 + words * i, *b = s2 + words * i;
        if ((int)(a[0] & 0xff) != n || (int)(b[0] & 0xff) != n)
            return fail(ctx, CSE_ERR_ARG, "pair %zu: header n does not match n=%d", i, n);
        if ((int)((a[0] >> 8) & 0xff) > n || (int)((b[0] >> 8) & 0xff) > n)
            return fail(ctx, CSE_ERR_ARG, "pair %zu: k > n", i);
    }
    HIPCHK(hipSetDevice(ctx->device));
    ENSURE(s1, words * 8 * npairs);
    ENSURE(s2, words * 8 * npairs);
    ENSURE(out, 12 * npairs);
    HIPCHK(hipMemcpyAsync(ctx->s1.p, s1, words * 8 * npairs, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->s2.p, s2, words * 8 * npairs, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->e0, ctx->stream));
    if (one_launch_per_pair) {
        for (size_t i = 0; i < npairs; i++) {
            hipLaunchKernelGGL(k_ip_pairs, dim3(1), dim3(64), 0, ctx->stream, dp<uint64_t>(ctx->s1) + words * i,
                               dp<uint64_t>(ctx->s2) + words * i, n, 1, dp<int32_t>(ctx->out) + 3 * i);
        }
    } else {
        const size_t blocks = (npairs + 3) / 4;
        hipLaunchKernelGGL(k_ip_pairs, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, dp<uint64_t>(ctx->s1),
                           dp<uint64_t>(ctx->s2), n, (int)npairs, dp<int32_t>(ctx->out));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->e1, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->out.p, 12 * npairs, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ms) *ms = elapsed(ctx->e0, ctx->e1);
    return CSE_OK;
}

int cse_prep_states(cse_ctx *ctx, int exact, int t, const uint64_t *L, int k, const uint64_t *idx, size_t count,
                    uint64_t *out) {
    if (!ctx) return CSE_ERR_ARG;
    if (t < 1 || t > CSE_NMAX) return fail(ctx, CSE_ERR_ARG, "t=%d outside [1, %d]", t, CSE_NMAX);
    if (count == 0) return CSE_OK;
    int rc = check_decomp(ctx, t, k, exact, L);
    if (rc) return rc;
    if (!idx || !out) return fail(ctx, CSE_ERR_ARG, "NULL buffer");
    if (count > kMaxStates) return fail(ctx, CSE_ERR_ARG, "%zu states above 2^30 per call", count);
    const uint64_t chi = exact ? (1ull << ((t + 1) / 2)) : (1ull << k);
    for (size_t c = 0; c < count; c++)
        if (idx[c] >= chi) return fail(ctx, CSE_ERR_ARG, "index %llu >= chi", (unsigned long long)idx[c]);
    HIPCHK(hipSetDevice(ctx->device));
    const size_t words = (size_t)CSE_STATE_WORDS(t);
    ENSURE(L, 8 * 64);
    ENSURE(idx, 8 * count);
    ENSURE(s1, 8 * words * count);
    if (!exact && k > 0) HIPCHK(hipMemcpyAsync(ctx->L.p, L, 8 * (size_t)k, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->idx.p, idx, 8 * count, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_prep_states, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, ctx->stream, exact, t,
                       dp<uint64_t>(ctx->L), k, dp<uint64_t>(ctx->idx), (int)count, dp<uint64_t>(ctx->s1));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->s1.p, 8 * words * count, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CSE_OK;
}

int cse_random_states(cse_ctx *ctx, int n, uint64_t seed, int64_t first, size_t count, const cse_projector *P,
                      uint64_t *out, int32_t *sexp) {
    if (!ctx) return CSE_ERR_ARG;
    if (n < 1 || n > CSE_NMAX) return fail(ctx, CSE_ERR_ARG, "n=%d outside [1, %d]", n, CSE_NMAX);
    if (count == 0) return CSE_OK;
    if (count > kMaxStates) return fail(ctx, CSE_ERR_ARG, "%zu states above 2^30 per call", count);
    if (P) {
        int rc = check_projector(ctx, P);
        if (rc) return rc;
        if (P->nqubits != n) return fail(ctx, CSE_ERR_ARG, "projector width %d != n %d", P->nqubits, n);
    }
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_thresholds(ctx);
    if (rc) return rc;
    const size_t words = (size_t)CSE_STATE_WORDS(n);
    ENSURE(s1, 8 * words * count);
    ENSURE(sexp, 4 * count);
    ENSURE(errflag, 8);
    rc = zero64(ctx, ctx->errflag.p, 1);
    if (rc) return rc;
    DevProjector dpj = to_dev(P);
    hipLaunchKernelGGL(k_random_states, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, ctx->stream, n,
                       dp<uint64_t>(ctx->thr) + 65 * n, seed, (long long)first, (int)count, dpj, P ? 1 : 0,
                       out ? dp<uint64_t>(ctx->s1) : nullptr, dp<int32_t>(ctx->sexp), dp<int>(ctx->errflag));
    HIPCHK(hipGetLastError());
    int errv = 0;
    HIPCHK(hipMemcpyAsync(&errv, ctx->errflag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (out) HIPCHK(hipMemcpyAsync(out, ctx->s1.p, 8 * words * count, hipMemcpyDeviceToHost, ctx->stream));
    if (sexp) HIPCHK(hipMemcpyAsync(sexp, ctx->sexp.p, 4 * count, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (errv) return fail(ctx, CSE_ERR_RNG, "random state draw exceeded its shrink cap");
    return CSE_OK;
}

// Sampled-mode instances [0, ninst): one record launch and one sweep over all
// of them, then each instance's estimator from its own samples.
// Runs job(w) for w in [0, nw): w >= 1 on threads of their own, w = 0 on the
// caller's.  A thread that cannot be started leaves its job to the caller.
extern "C++" template <class F>
static void run_jobs(int nw, F &&job) {
    std::vector<std::thread> pool;
    int w = 1;
    try {
        for (; w < nw; w++) pool.emplace_back(job, w);
    } catch (...) {
    }
    for (int v = w; v < nw; v++) job(v);
    job(0);
    for (std::thread &th : pool) th.join();
}

// One instance's estimators over count samples: fixed parts of kEstPart
// samples (so the Kahan mean depends on count alone), spread over nthr host
// threads.  At 2^16 samples one thread takes about 0.5 ms.
constexpr int64_t kEstPart = 8192;
static void split_estimators(int t, int64_t count, const long long *z, const int32_t *se, const double *kz,
                             int nthr, double *mean, double *mean_kahan) {
    const int nparts = (int)((count + kEstPart - 1) / kEstPart);
    if (nparts < 2) {
        cse_sampled_estimators(t, count, z, se, kz, mean, mean_kahan);
        return;
    }
    std::vector<cse_est_part> parts((size_t)nparts);
    nthr = std::max(1, std::min(nthr, nparts));
    auto work = [&](int w) {
        for (int q = w; q < nparts; q += nthr)
            cse_est_accumulate(&parts[(size_t)q], t, q * kEstPart, std::min(count, (q + 1) * kEstPart), z, se, kz);
    };
    run_jobs(nthr, work);
    cse_est_finish(parts.data(), nparts, t, count, mean, mean_kahan);
}

static int sampled_batch(cse_ctx *ctx, const cse_sampled_instance *in, int ninst, int chunk,
                         cse_sampled_result *res, double *kernel_ms) {
    PlanStore store(ninst);
    Plan pl = store.plan();
    bool want_terms = false;
    for (int i = 0; i < ninst; i++) {
        InstDesc d = make_inst(in[i].P.nqubits, in[i].k, in[i].exact, in[i].L, chunk);
        d.seed = in[i].seed;
        d.first = in[i].first;
        pl.add(d, to_dev(&in[i].P), in[i].count, in[i].count * d.nchunks, in[i].count * d.chi);
        want_terms = want_terms || res[i].terms != nullptr;
    }
    if (pl.nrec > kMaxSamples) return fail(ctx, CSE_ERR_ARG, "%lld samples above 2^24 per call", pl.nrec);
    if (want_terms && pl.nterms > (1ll << 31)) return fail(ctx, CSE_ERR_ARG, "per-term output too large");
    int rc = ensure_thresholds(ctx);
    if (!rc) rc = bind_ctx_mem(ctx, pl, 0, want_terms);
    if (rc) return rc;
    bool graphed = false;
    rc = run_batch(ctx, pl, 0, want_terms, stage_plan(ctx, pl), &graphed);
    if (rc) return rc;

    // the results block came back in one copy, into a pinned stage of the
    // same layout
    const OutLayout o = out_layout(pl, 0);
    char *hs = ctx->b_hstage;
    const long long *z = (const long long *)(hs + o.zsum);
    const double *kz = (const double *)(hs + o.kahan);
    const int32_t *se = (const int32_t *)(hs + o.sexp);
    const unsigned long long *nips = (const unsigned long long *)(hs + o.ips);
    int *errp = (int *)(hs + o.err);
    for (int i = 0; i < ninst; i++)
        if (res[i].terms)
            HIPCHK(hipMemcpyAsync(res[i].terms, ctx->b_terms + pl.inst[i].term_off,
                                  4 * (size_t)(in[i].count * pl.inst[i].chi), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int errv = *errp;
    if (errv & 1) return fail(ctx, CSE_ERR_RNG, "random state draw exceeded its shrink cap");
    if (errv & 2) return fail(ctx, CSE_ERR_HIP, "internal: an inner product exponent left [-64, 0]");
    // a replayed graph is timed as a whole (records and sweep together)
    const double pms = graphed ? 0.0 : elapsed(ctx->e0, ctx->e1);
    const double kms = graphed ? elapsed(ctx->e0, ctx->e2) : elapsed(ctx->e1, ctx->e2);
    if (kernel_ms) *kernel_ms = kms;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthr = (ninst >= 8 && pl.nrec >= (1 << 16)) ? (int)std::min(8u, hw ? hw : 1u) : 1;
    // with one thread per group of instances, each instance's samples take
    // one thread; otherwise a large instance splits its samples
    const int sthr = nthr > 1 ? 1 : (int)std::min(8u, hw ? hw : 1u);
    auto convert = [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            cse_sampled_result *r = &res[i];
            const size_t off = (size_t)pl.rec_off[i], cnt = (size_t)in[i].count;
            r->prep_ms = pms;
            r->kernel_ms = kms;
            r->ips = (int64_t)nips[i];
            r->vanished = 0;
            for (size_t s = 0; s < cnt; s++)
                if (se[off + s] < 0) r->vanished++;
            split_estimators(in[i].P.nqubits, (int64_t)cnt, z + 4 * off, se + off, kz + 2 * off,
                             cnt >= 4 * (size_t)kEstPart ? sthr : 1, &r->mean, &r->mean_kahan);
            if (r->zsum) memcpy(r->zsum, z + 4 * off, 32 * cnt);
            if (r->sexp) memcpy(r->sexp, se + off, 4 * cnt);
            if (r->kahan) memcpy(r->kahan, kz + 2 * off, 16 * cnt);
        }
    };
    // the estimators cost about as much per sample as the copies: split large
    // batches of instances over host threads
    if (nthr > 1) {
        const int per = (ninst + nthr - 1) / nthr;
        run_jobs(nthr, [&](int w) { convert(std::min(ninst, w * per), std::min(ninst, (w + 1) * per)); });
    } else {
        convert(0, ninst);
    }
    return CSE_OK;
}

int cse_sampled_projector(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k, int exact,
                          uint64_t seed, int64_t first, int64_t count, int chunk, cse_sampled_result *res) {
    if (!ctx || !res) return CSE_ERR_ARG;
    int rc = check_projector(ctx, P);
    if (rc) return rc;
    rc = check_decomp(ctx, P->nqubits, k, exact, L);
    if (rc) return rc;
    if (count < 1) return fail(ctx, CSE_ERR_ARG, "count=%lld", (long long)count);
    if (count > kMaxSamples) return fail(ctx, CSE_ERR_ARG, "count=%lld above 2^24 per call", (long long)count);
    HIPCHK(hipSetDevice(ctx->device));
    cse_sampled_instance in;
    memset(&in, 0, sizeof(in));
    in.P = *P;
    in.exact = exact;
    in.k = exact ? 0 : k;
    if (!exact)
        for (int r = 0; r < k; r++) in.L[r] = L[r];
    in.seed = seed;
    in.first = first;
    in.count = count;
    return sampled_batch(ctx, &in, 1, chunk, res, nullptr);
}

int cse_batch_sampled(cse_ctx *ctx, const cse_sampled_instance *inst, int ninst, int chunk,
                      cse_sampled_result *res, double *kernel_ms) {
    if (!ctx || !inst || !res || ninst < 1) return CSE_ERR_ARG;
    long long total = 0;
    for (int i = 0; i < ninst; i++) {
        int rc = check_projector(ctx, &inst[i].P);
        if (!rc) rc = check_decomp(ctx, inst[i].P.nqubits, inst[i].k, inst[i].exact, inst[i].L);
        if (rc) return fail(ctx, rc, "instance %d: %s", i, ctx->err);
        if (inst[i].count < 1) return fail(ctx, CSE_ERR_ARG, "instance %d: count=%lld", i, (long long)inst[i].count);
        total += inst[i].count;
        if (total > kMaxSamples) return fail(ctx, CSE_ERR_ARG, "more than 2^24 samples in one call");
    }
    HIPCHK(hipSetDevice(ctx->device));
    return sampled_batch(ctx, inst, ninst, chunk, res, kernel_ms);
}

// Exact-mode instances [0, ninst) in one record launch and one sweep.  With
// use_static the call works in this device's static scratch array and the
// host stages hup / hstage (of at least kStaticPlanBytes / kStaticOutBytes)
// instead of the context's buffers, and allocates nothing.
static int exact_batch(cse_ctx *ctx, const cse_instance *in, int ninst, int chunk, cse_exact_result *res,
                       int32_t *terms_out, double *kernel_ms, bool use_static = false, char *hup = nullptr,
                       char *hstage = nullptr) {
    if (chunk <= 0) chunk = exact_chunk(ctx, in, ninst);
    // the static path keeps its one-instance plan on the stack
    InstDesc s_inst[1];
    long long s_ro[2], s_io[2];
    DevProjector s_pj[1];
    PlanStore store(use_static ? 0 : ninst);
    Plan pl = store.plan();
    if (use_static) {
        if (ninst != 1) return fail(ctx, CSE_ERR_ARG, "internal: static call with %d instances", ninst);
        pl.inst = s_inst;
        pl.rec_off = s_ro;
        pl.item_off = s_io;
        pl.projs = s_pj;
    }
    // items are slices of CSE_SLICE_J j; a claim takes at most chunk / CSE_SLICE_J of them
    pl.take_max = chunk / CSE_SLICE_J > 1 ? chunk / CSE_SLICE_J : 1;
    for (int i = 0; i < ninst; i++) {
        const InstDesc d = make_inst(in[i].P.nqubits, in[i].k, in[i].exact, in[i].L, CSE_SLICE_J);
        pl.add(d, to_dev(&in[i].P), d.chi, tri_items(d.chi, d.chunk, d.nchunks), d.chi * (d.chi + 1) / 2);
    }
    if (terms_out && pl.nterms > (1ll << 31)) return fail(ctx, CSE_ERR_ARG, "per-term output too large");
    int rc = use_static ? bind_static_mem(ctx, pl, hup, hstage) : bind_ctx_mem(ctx, pl, 1, terms_out != nullptr);
    if (rc) return rc;
    // the static path never captures (graph memory nodes would be
    // allocations of their own)
    bool graphed = false;
    const size_t plan_bytes = stage_plan(ctx, pl);
    rc = use_static ? enqueue_batch(ctx, pl, 1, false, plan_bytes, true)
                    : run_batch(ctx, pl, 1, terms_out != nullptr, plan_bytes, &graphed);
    if (rc) return rc;
    // the results block came back in one copy, into a pinned stage of the
    // same layout: histograms (HIST_STRIDE words per instance), the Kahan
    // pairs, the inner-product counts, the error flag
    const OutLayout o = out_layout(pl, 1);
    char *hs = ctx->b_hstage;
    const unsigned long long *hist = (const unsigned long long *)(hs + o.hist);
    const double *kz = (const double *)(hs + o.kahan);
    const unsigned long long *nips = (const unsigned long long *)(hs + o.ips);
    int *errp = (int *)(hs + o.err);
    if (terms_out)
        HIPCHK(hipMemcpyAsync(terms_out, ctx->b_terms, 4 * (size_t)pl.nterms, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (*errp & 2) return fail(ctx, CSE_ERR_HIP, "internal: an inner product exponent left [-64, 0]");
    const double tms = elapsed(ctx->e0, ctx->e2);
    const double kms = graphed ? tms : elapsed(ctx->e1, ctx->e2);
    if (kernel_ms) *kernel_ms = kms;
    // each instance's exact value from its histogram (a few microseconds
    // each): split over host threads for large batches
    auto convert = [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            cse_exact_result *r = &res[i];
            const int64_t *h = (const int64_t *)&hist[(size_t)HIST_STRIDE * i];
            r->pmin = CSE_HIST_PMIN;
            if (r->hist) memcpy(r->hist, h, 8 * (size_t)HIST_STRIDE);
            cse_hist_value(h, CSE_HIST_PMIN, &r->re, &r->im);
            r->value = hypot(r->re, r->im);
            r->kahan_re = kz[2 * i];
            r->kahan_im = kz[2 * i + 1];
            r->terms = pl.inst[i].chi * (pl.inst[i].chi + 1) / 2;
            r->ips = (int64_t)nips[i];
            r->kernel_ms = kms;
            r->total_ms = tms;
        }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthr = ninst >= 512 ? (int)std::min(8u, hw ? hw : 1u) : 1;
    if (nthr > 1) {
        const int per = (ninst + nthr - 1) / nthr;
        run_jobs(nthr, [&](int w) { convert(std::min(ninst, w * per), std::min(ninst, (w + 1) * per)); });
    } else {
        convert(0, ninst);
    }
    return CSE_OK;
}

int cse_exact_projector(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k, int exact,
                        cse_exact_result *res) {
    if (!ctx || !res) return CSE_ERR_ARG;
    int rc = check_projector(ctx, P);
    if (rc) return rc;
    rc = check_decomp(ctx, P->nqubits, k, exact, L);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    cse_instance in;
    memset(&in, 0, sizeof(in));
    in.P = *P;
    in.exact = exact;
    in.k = exact ? 0 : k;
    if (!exact)
        for (int r = 0; r < k; r++) in.L[r] = L[r];
    return exact_batch(ctx, &in, 1, 0, res, nullptr, nullptr);
}

int cse_exact_projector_noalloc(cse_ctx *ctx, const cse_projector *P, const uint64_t *L, int k, int exact,
                                cse_exact_result *res) {
    if (!ctx || !res) return CSE_ERR_ARG;
    int rc = check_projector(ctx, P);
    if (rc) return rc;
    rc = check_decomp(ctx, P->nqubits, k, exact, L);
    if (rc) return rc;
    const int logchi = exact ? (P->nqubits + 1) / 2 : k;
    if (logchi > kStaticMaxLogChi)
        return fail(ctx, CSE_ERR_ARG, "chi=2^%d above the no-allocation limit 2^%d", logchi, kStaticMaxLogChi);
    HIPCHK(hipSetDevice(ctx->device));
    cse_instance in;
    memset(&in, 0, sizeof(in));
    in.P = *P;
    in.exact = exact;
    in.k = exact ? 0 : k;
    if (!exact)
        for (int r = 0; r < k; r++) in.L[r] = L[r];
    // host stages on the stack: pageable copies, synchronized before return
    alignas(256) char hup[kStaticPlanBytes];
    alignas(256) char hstage[kStaticOutBytes];
    return exact_batch(ctx, &in, 1, 0, res, nullptr, nullptr, true, hup, hstage);
}

int cse_batch_exact(cse_ctx *ctx, const cse_instance *inst, int ninst, int chunk, cse_exact_result *res,
                    int32_t *terms_out, double *kernel_ms) {
    if (!ctx || !inst || !res || ninst < 1) return CSE_ERR_ARG;
    for (int i = 0; i < ninst; i++) {
        int rc = check_projector(ctx, &inst[i].P);
        if (!rc) rc = check_decomp(ctx, inst[i].P.nqubits, inst[i].k, inst[i].exact, inst[i].L);
        if (rc) return fail(ctx, rc, "instance %d: %s", i, ctx->err);
    }
    HIPCHK(hipSetDevice(ctx->device));
    return exact_batch(ctx, inst, ninst, chunk, res, terms_out, kernel_ms);
}

int cse_lspace_weights(cse_ctx *ctx, const uint64_t *L, int k, int64_t *hist) {
    if (!ctx || !hist || (k > 0 && !L)) return CSE_ERR_ARG;
    if (k < 0 || k > 40) return fail(ctx, CSE_ERR_ARG, "k=%d outside [0, 40]", k);
    HIPCHK(hipSetDevice(ctx->device));
    ENSURE(lsh, 8 * 65);
    ENSURE(L, 8 * (size_t)CSE_NMAX);
    if (k > 0) HIPCHK(hipMemcpyAsync(ctx->L.p, L, 8 * (size_t)k, hipMemcpyHostToDevice, ctx->stream));
    int rc = zero64(ctx, ctx->lsh.p, 65);
    if (rc) return rc;
    // 2^(k - shift) threads, each walking 2^shift combinations
    const int shift = k > 18 ? k - 18 : 0;
    const uint64_t threads = 1ull << (k - shift);
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    hipLaunchKernelGGL(k_lspace_hist, dim3(blocks), dim3(256), 0, ctx->stream, dp<uint64_t>(ctx->L), k, shift,
                       dp<unsigned long long>(ctx->lsh));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hist, ctx->lsh.p, 8 * 65, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CSE_OK;
}

}  // extern "C"
