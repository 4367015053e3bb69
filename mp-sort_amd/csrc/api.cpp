// api.cpp -- C ABI of the MI355X histogram sort: argument checks, device
// workspaces, the histogram-sort pipeline (mpsort-mpi.c:161-604 with P virtual
// ranks), timers, and the reference-compatible entry points of mpsort.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "internal.h"
#include "../../include/mpsort_gpu.h"

namespace mpsg {

// histsort.hip
int launch_bisect(const void* recs, int rw, const KeySpec& ks, const int64_t* sbeg,
                  const int64_t* scnt, int P, const int64_t* C, int64_t* win, int64_t* myclt,
                  int64_t* mycle, uint64_t* split, uint64_t* pminmax, int* rounds, uint32_t* err,
                  hipStream_t st);
int launch_layout(int P, const int64_t* C, const int64_t* scnt, const int64_t* sbeg, int64_t* myclt,
                  int64_t* mycle, int64_t* myC, int64_t* send, int64_t* run_dst, int64_t* run_src,
                  uint32_t* err, hipStream_t st);
int launch_exchange(const void* src, void* dst, int64_t ntot, int rw, int nruns,
                    const int64_t* run_dst, const int64_t* run_src, uint32_t* err, hipStream_t st);
int launch_exchange_hist(const void* src, void* dst, int64_t ntot, int rw, int nruns,
                         const int64_t* run_dst, const int64_t* run_src, uint32_t* err,
                         const int64_t* C, int P, const KeySpec& ks, uint32_t* hist, hipStream_t st);
int launch_pack(const void* in, int64_t n, int64_t elsize, int64_t key_off, int key_bytes,
                void* out, int rw, long long* checksum, hipStream_t st);
int launch_gather(const void* in, int64_t n, int64_t elsize, const void* srec, int rw, void* out,
                  long long* checksum, uint32_t* err, hipStream_t st);

static char g_err[1024];
static int g_options = 0;
static int g_vranks = 0;  // 0 = not set yet (environment decides)
static mpsort_gpu_malloc_fn g_malloc = nullptr;
static mpsort_gpu_free_fn g_free = nullptr;
static void* g_alloc_ud = nullptr;
static mpsort_gpu_stats g_last;

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int set_hip_error(hipError_t e, const char* what, const char* file, int line) {
    snprintf(g_err, sizeof g_err, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
             what, file, line);
    return MPSORT_ERR_HIP;
}

static void* dev_alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (g_malloc) return g_malloc(bytes, g_alloc_ud);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}
static void dev_free(void* p) {
    if (!p) return;
    if (g_free) g_free(p, g_alloc_ud);
    else (void)hipFree(p);
}

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    bool zero_on_grow = false;
    int ensure(size_t need) {
        if (need <= bytes && p) return 0;
        dev_free(p);
        p = nullptr;
        bytes = 0;
        size_t want = need + need / 8 + 4096;
        p = dev_alloc(want);
        if (!p) return fail(MPSORT_ERR_NOMEM, "device allocation of %zu bytes failed", want);
        bytes = want;
        if (zero_on_grow) {
            if (hipMemset(p, 0, want) != hipSuccess) return fail(MPSORT_ERR_HIP, "hipMemset failed");
        }
        return 0;
    }
    void release() {
        dev_free(p);
        p = nullptr;
        bytes = 0;
    }
};

enum { kEvStart, kEvFirst, kEvFindP, kEvLayout, kEvExchange, kEvSecond, kEvEnd, kEvPhase };

struct DevCtx {
    bool init = false;
    int device = 0;
    hipStream_t st = nullptr;
    uint32_t epoch = 0;
    Buf din, dout;      // host staging
    Buf s0, s1, rbuf;   // sort-record buffers
    Buf small;          // counts, tables
    Buf lookback;       // look-back flags (only ever holds flag words)
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t ev[kEvPhase];
    hipEvent_t pass_ev[2][2 * kMaxPasses];
    // last call's device arrays for mpsort_gpu_last_counts
    int last_P = 0;
    int64_t* C = nullptr;
    int64_t *myclt = nullptr, *mycle = nullptr, *myC = nullptr, *send = nullptr;
    uint64_t *split = nullptr, *pminmax = nullptr;
    int last_width = 8, last_nwords = 1, last_signed = 0, last_callback = 0;
    int64_t last_rsize = 8;
    int64_t last_ntot = 0;
};
static DevCtx g_ctx[64];

uint32_t next_epoch(int device) {
    DevCtx& c = g_ctx[device & 63];
    c.epoch++;
    if (c.epoch >= (1u << 23)) {
        // flags of old launches could alias new epochs: clear and restart
        (void)hipStreamSynchronize(c.st);
        if (c.lookback.p) (void)hipMemset(c.lookback.p, 0, c.lookback.bytes);
        c.epoch = 1;
    }
    return c.epoch;
}

static int get_ctx(DevCtx** out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(MPSORT_ERR_NODEVICE, "no GPU visible to the HIP runtime");
    int dev = 0;
    MPSG_CHECK(hipGetDevice(&dev));
    DevCtx& c = g_ctx[dev & 63];
    if (!c.init) {
        c.device = dev;
        MPSG_CHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
        for (int i = 0; i < kEvPhase; i++) MPSG_CHECK(hipEventCreate(&c.ev[i]));
        for (int a = 0; a < 2; a++)
            for (int i = 0; i < 2 * kMaxPasses; i++) MPSG_CHECK(hipEventCreate(&c.pass_ev[a][i]));
        c.lookback.zero_on_grow = true;
        c.init = true;
    }
    *out = &c;
    return 0;
}

static int ensure_pinned(DevCtx& c, size_t bytes) {
    if (bytes <= c.pinned_bytes) return 0;
    if (c.pinned) (void)hipHostFree(c.pinned);
    c.pinned = nullptr;
    c.pinned_bytes = 0;
    size_t want = bytes + 4096;
    MPSG_CHECK(hipHostMalloc(&c.pinned, want, hipHostMallocDefault));
    c.pinned_bytes = want;
    return 0;
}

// bump allocator over one device buffer
struct Arena {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = reinterpret_cast<T*>(base + off);
        off += n * sizeof(T);
        return p;
    }
};
struct Sizer {
    size_t off = 0;
    template <typename T>
    void take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        off += n * sizeof(T);
    }
};

struct PlanArrays {
    int64_t *begin, *count, *tile_start;
    uint32_t *hist, *pass_flags, *tickets;
    int64_t* digit_base;
};

template <class A>
static void lay_plan(A& a, int nseg, int npass, PlanArrays* pa) {
    if constexpr (std::is_same<A, Arena>::value) {
        pa->begin = a.template take<int64_t>(nseg);
        pa->count = a.template take<int64_t>(nseg);
        pa->tile_start = a.template take<int64_t>(nseg + 1);
        pa->hist = a.template take<uint32_t>((size_t)nseg * npass * kBins);
        pa->digit_base = a.template take<int64_t>((size_t)nseg * npass * kBins);
        pa->pass_flags = a.template take<uint32_t>(kMaxPasses + 1);
        pa->tickets = a.template take<uint32_t>(kMaxPasses);
    } else {
        a.template take<int64_t>(nseg);
        a.template take<int64_t>(nseg);
        a.template take<int64_t>(nseg + 1);
        a.template take<uint32_t>((size_t)nseg * npass * kBins);
        a.template take<int64_t>((size_t)nseg * npass * kBins);
        a.template take<uint32_t>(kMaxPasses + 1);
        a.template take<uint32_t>(kMaxPasses);
    }
}

struct HistArrays {
    int64_t *C, *win, *myclt, *mycle, *myC, *send, *run_dst, *run_src;
    uint64_t *split, *pminmax;
    int* rounds;
    uint32_t* err;
    long long* checksum;
};

template <class A>
static void lay_hist(A& a, int P, HistArrays* h) {
    const size_t P1 = (size_t)P + 1, PP = (size_t)P * P, Pm = P > 1 ? (size_t)P - 1 : 1;
    if constexpr (std::is_same<A, Arena>::value) {
        h->C = a.template take<int64_t>(P1);
        h->win = a.template take<int64_t>(Pm * P * 4);
        h->myclt = a.template take<int64_t>(P * P1);
        h->mycle = a.template take<int64_t>(P * P1);
        h->myC = a.template take<int64_t>(P * P1);
        h->send = a.template take<int64_t>(PP);
        h->run_dst = a.template take<int64_t>(PP);
        h->run_src = a.template take<int64_t>(PP);
        h->split = a.template take<uint64_t>(Pm * 4);
        h->pminmax = a.template take<uint64_t>(8);
        h->rounds = a.template take<int>(Pm);
        h->err = a.template take<uint32_t>(4);
        h->checksum = a.template take<long long>(2);
    } else {
        a.template take<int64_t>(P1);
        a.template take<int64_t>(Pm * P * 4);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(PP);
        a.template take<int64_t>(PP);
        a.template take<int64_t>(PP);
        a.template take<uint64_t>(Pm * 4);
        a.template take<uint64_t>(8);
        a.template take<int>(Pm);
        a.template take<uint32_t>(4);
        a.template take<long long>(2);
    }
}

static void fill_keyspec(KeySpec* ks, int rw, int key_off, int key_bytes, int width, int is_signed) {
    memset(ks, 0, sizeof *ks);
    ks->rec_words = rw;
    ks->key_off = key_off;
    ks->key_bytes = key_bytes;
    if (is_signed && width > 0)
        for (int b = width - 1; b < key_bytes; b += width) ks->xmask[b] = 0x80;
}

// host description of one segmented sort: segment begins/counts -> tiles
static int64_t make_tiles(const int64_t* counts, int nseg, int64_t T, int64_t* begin, int64_t* tile_start) {
    int64_t b = 0, t = 0;
    for (int s = 0; s < nseg; s++) {
        begin[s] = b;
        tile_start[s] = t;
        b += counts[s];
        t += (counts[s] + T - 1) / T;
    }
    tile_start[nseg] = t;
    return t;
}

struct CallShape {
    int P;
    int rw;
    KeySpec ks;
    int64_t ntot;
    bool general;
    int64_t elsize;
    int64_t key_off_user;  // general path: key bytes at this offset of the caller's record
    int key_bytes_user;
    bool checksum;
    bool timing;
};

static double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

// The histogram sort on device-resident data.  d_in / d_out are the caller's
// records (device); for the general path they are packed into sort records.
static int run_histsort(DevCtx& c, const CallShape& cs, const void* d_in, void* d_out,
                        const int64_t* nin, const int64_t* nout, const void* packed_host,
                        mpsort_gpu_stats* stats) {
    hipStream_t st = c.st;
    const int P = cs.P;
    const int rw = cs.rw;
    const int64_t n = cs.ntot;
    const size_t recbytes = (size_t)rw * 4;
    const int npass = cs.ks.key_bytes;

    // ---- workspace
    Sizer sz;
    lay_plan(sz, P, npass, nullptr);
    lay_plan(sz, P, npass, nullptr);
    lay_hist(sz, P, nullptr);
    int rc = c.small.ensure(sz.off + 4096);
    if (rc) return rc;
    Arena ar{(char*)c.small.p};
    PlanArrays pa, pb;
    HistArrays h;
    lay_plan(ar, P, npass, &pa);
    lay_plan(ar, P, npass, &pb);
    lay_hist(ar, P, &h);

    if (P > 1 || cs.general) {
        if ((rc = c.s0.ensure(n * recbytes))) return rc;
    }
    if ((rc = c.s1.ensure(n * recbytes))) return rc;
    if (cs.general && (rc = c.rbuf.ensure(n * recbytes))) return rc;

    // host tables: [beginA, countA, tsA(P+1), beginB, countB, tsB(P+1), C(P+1)]
    const size_t tab = (size_t)(7 * P + 4);
    if ((rc = ensure_pinned(c, tab * sizeof(int64_t) + 4096))) return rc;
    int64_t* ht = (int64_t*)c.pinned;
    int64_t *hbA = ht, *hcA = ht + P, *htA = ht + 2 * P;
    int64_t *hbB = ht + 3 * P + 1, *hcB = ht + 4 * P + 1, *htB = ht + 5 * P + 1;
    int64_t* hC = ht + 6 * P + 2;
    for (int t = 0; t < P; t++) { hcA[t] = nin[t]; hcB[t] = nout[t]; }
    const TileCfg tile = default_tile(rw);
    const int64_t trecs = (int64_t)tile.block * tile.items;
    const int64_t ntA = make_tiles(hcA, P, trecs, hbA, htA);
    const int64_t ntB = make_tiles(hcB, P, trecs, hbB, htB);
    hC[0] = 0;
    for (int t = 0; t < P; t++) hC[t + 1] = hC[t] + nout[t];
    const int64_t ntiles = ntA > ntB ? ntA : ntB;
    if ((rc = c.lookback.ensure((size_t)(ntiles > 0 ? ntiles : 1) * kBins * sizeof(uint64_t)))) return rc;

    MPSG_CHECK(hipMemcpyAsync(pa.begin, hbA, sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pa.count, hcA, sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pa.tile_start, htA, sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pb.begin, hbB, sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pb.count, hcB, sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pb.tile_start, htB, sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(h.C, hC, sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
    if ((rc = launch_fill_u32(h.err, 0u, 4, st))) return rc;
    if ((rc = launch_fill_u32((uint32_t*)h.checksum, 0u, 4, st))) return rc;
    // both plans' flag words (the second plan is unused when P == 1)
    if ((rc = launch_fill_u32(pa.pass_flags, 0u, kMaxPasses + 1, st))) return rc;
    if ((rc = launch_fill_u32(pb.pass_flags, 0u, kMaxPasses + 1, st))) return rc;

    auto mkplan = [&](const PlanArrays& a, int64_t nt) {
        SegSortPlan p;
        p.rw = rw;
        p.tile = tile;
        p.ks = cs.ks;
        p.nrec = n;
        p.nseg = P;
        p.ntiles = nt;
        p.d_begin = a.begin;
        p.d_count = a.count;
        p.d_tile_start = a.tile_start;
        p.d_hist = a.hist;
        p.d_digit_base = a.digit_base;
        p.d_pass_flags = a.pass_flags;
        p.d_lookback = (uint64_t*)c.lookback.p;
        p.d_tickets = a.tickets;
        p.d_checksum = h.checksum;
        return p;
    };
    const SegSortPlan planA = mkplan(pa, ntA);
    const SegSortPlan planB = mkplan(pb, ntB);

    MPSG_CHECK(hipEventRecord(c.ev[kEvStart], st));

    // general path: pack {key, index} sort records
    void* recs_in = const_cast<void*>(d_in);
    void* final_dst = d_out;
    if (cs.general) {
        if (packed_host) {
            // the radix-callback form: sort records were built on the host
            MPSG_CHECK(hipMemcpyAsync(c.rbuf.p, packed_host, n * recbytes, hipMemcpyHostToDevice, st));
        } else {
            if ((rc = launch_pack(d_in, n, cs.elsize, cs.key_off_user, cs.key_bytes_user, c.rbuf.p, rw,
                                  cs.checksum ? h.checksum : nullptr, st)))
                return rc;
        }
        recs_in = c.rbuf.p;
        final_dst = c.rbuf.p;
    }
    const bool cs_direct = cs.checksum && !cs.general;

    int first_passes = 0, second_passes = 0;
    hipEvent_t* pevA = cs.timing ? c.pass_ev[0] : nullptr;
    hipEvent_t* pevB = cs.timing ? c.pass_ev[1] : nullptr;
    if (P == 1) {
        // one rank: FirstSort is the whole sort (the exchange is the identity
        // and SecondSort re-sorts a sorted run)
        SortBuffers b{recs_in, final_dst, c.s1.p};
        if ((rc = segmented_sort(planA, b, st, cs_direct, cs_direct, &first_passes, pevA))) return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        for (int e = kEvFindP; e <= kEvSecond; e++) MPSG_CHECK(hipEventRecord(c.ev[e], st));
    } else {
        SortBuffers b1{recs_in, c.s0.p, c.s1.p};
        if ((rc = segmented_sort(planA, b1, st, cs_direct, false, &first_passes, pevA))) return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        if ((rc = launch_bisect(c.s0.p, rw, cs.ks, pa.begin, pa.count, P, h.C, h.win, h.myclt, h.mycle,
                                h.split, h.pminmax, h.rounds, h.err, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFindP], st));
        if ((rc = launch_layout(P, h.C, pa.count, pa.begin, h.myclt, h.mycle, h.myC, h.send, h.run_dst,
                                h.run_src, h.err, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvLayout], st));
        // the exchange also builds the SecondSort's digit histograms
        if ((rc = seg_clear(planB, st))) return rc;
        if ((rc = launch_exchange_hist(c.s0.p, c.s1.p, n, rw, P * P, h.run_dst, h.run_src, h.err, h.C, P,
                                       cs.ks, planB.d_hist, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvExchange], st));
        SortBuffers b2{c.s1.p, final_dst, c.s0.p};
        if ((rc = seg_passes(planB, b2, st, cs_direct ? h.checksum + 1 : nullptr, &second_passes, pevB)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvSecond], st));
    }
    if (cs.general) {
        if ((rc = launch_gather(d_in, n, cs.elsize, c.rbuf.p, rw, d_out,
                                cs.checksum ? h.checksum + 1 : nullptr, h.err, st)))
            return rc;
    }
    MPSG_CHECK(hipEventRecord(c.ev[kEvEnd], st));

    // ---- results
    uint32_t errs[3] = {0, 0, 0};
    long long sums[2] = {0, 0};
    int rounds_host = 0;
    MPSG_CHECK(hipMemcpyAsync(&errs[0], h.err, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipMemcpyAsync(&errs[1], pa.pass_flags + kMaxPasses, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipMemcpyAsync(&errs[2], pb.pass_flags + kMaxPasses, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipMemcpyAsync(sums, h.checksum, sizeof sums, hipMemcpyDeviceToHost, st));
    std::vector<int> rounds(P > 1 ? P - 1 : 1, 0);
    if (P > 1)
        MPSG_CHECK(hipMemcpyAsync(rounds.data(), h.rounds, sizeof(int) * (P - 1), hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i + 1 < P; i++) rounds_host = rounds[i] > rounds_host ? rounds[i] : rounds_host;

    // stats
    stats->nvranks = P;
    stats->record_bytes = (int32_t)recbytes;
    stats->first_passes = first_passes;
    stats->second_passes = second_passes;
    stats->bisect_rounds = rounds_host;
    stats->general_path = cs.general ? 1 : 0;
    stats->nrecords = n;
    stats->checksum_in = sums[0];
    stats->checksum_out = sums[1];
    stats->t_first_sort_ms = ev_ms(c.ev[kEvStart], c.ev[kEvFirst]);
    stats->t_findp_ms = ev_ms(c.ev[kEvFirst], c.ev[kEvFindP]);
    stats->t_laysolve_ms = ev_ms(c.ev[kEvFindP], c.ev[kEvLayout]);
    stats->t_exchange_ms = ev_ms(c.ev[kEvLayout], c.ev[kEvExchange]);
    stats->t_second_sort_ms = ev_ms(c.ev[kEvExchange], c.ev[kEvSecond]);
    stats->t_total_ms = ev_ms(c.ev[kEvStart], c.ev[kEvEnd]);
    int np = 0;
    if (cs.timing) {
        for (int i = 0; i < first_passes && np < 64; i++)
            stats->pass_ms[np++] = ev_ms(c.pass_ev[0][2 * i], c.pass_ev[0][2 * i + 1]);
        for (int i = 0; i < second_passes && np < 64; i++)
            stats->pass_ms[np++] = ev_ms(c.pass_ev[1][2 * i], c.pass_ev[1][2 * i + 1]);
    }
    stats->npass_ms = np;

    c.last_P = P;
    c.C = h.C;
    c.myclt = h.myclt;
    c.mycle = h.mycle;
    c.myC = h.myC;
    c.send = h.send;
    c.split = h.split;
    c.pminmax = h.pminmax;
    c.last_ntot = n;

    if (errs[1] | errs[2] | (errs[0] & 1u))
        return fail(MPSORT_ERR_TIMEOUT, "a look-back wait exceeded its spin limit (flags %u %u %u)", errs[0],
                    errs[1], errs[2]);
    if (errs[0])
        return fail(MPSORT_ERR_LAYOUT, "layout check failed (device flags 0x%x): the splitter search "
                                       "or layout solve is inconsistent",
                    errs[0]);
    if (cs.checksum && sums[0] != sums[1])
        return fail(MPSORT_ERR_CHECKSUM, "MPSort: Data changed after sorting; checksum mismatch (%lld != %lld)",
                    sums[0], sums[1]);
    return 0;
}

static int validate_counts(const int64_t* nin, int P, const int64_t* nout, int64_t* ntot) {
    if (P < 1 || P > 4096) return fail(MPSORT_ERR_ARG, "number of virtual ranks %d not in [1, 4096]", P);
    if (!nin || !nout) return fail(MPSORT_ERR_ARG, "nin/nout must not be NULL");
    int64_t a = 0, b = 0;
    for (int t = 0; t < P; t++) {
        if (nin[t] < 0 || nout[t] < 0) return fail(MPSORT_ERR_ARG, "negative count for rank %d", t);
        a += nin[t];
        b += nout[t];
    }
    if (a != b)
        return fail(MPSORT_ERR_SIZE_MISMATCH,
                    "MPSort: total number of items in the item does not match the input %lld != %lld", (long long)b,
                    (long long)a);
    *ntot = a;
    return 0;
}

static int shape_for(const mpsort_key_desc* key, size_t elsize, unsigned flags, const void* base, const void* out,
                     CallShape* cs) {
    if (!key) return fail(MPSORT_ERR_ARG, "key descriptor is NULL");
    if (elsize == 0) return fail(MPSORT_ERR_ARG, "elsize is 0");
    if (key->width != 4 && key->width != 8) return fail(MPSORT_ERR_UNSUPPORTED, "key word width %d not 4 or 8", key->width);
    if (key->nwords < 1) return fail(MPSORT_ERR_ARG, "key nwords %d < 1", key->nwords);
    const int64_t kb = (int64_t)key->width * key->nwords;
    if (key->offset < 0 || key->offset + kb > (int64_t)elsize)
        return fail(MPSORT_ERR_ARG, "key [%lld, %lld) outside the %zu-byte record", (long long)key->offset,
                    (long long)(key->offset + kb), elsize);
    if (kb > kMaxKeyBytes) return fail(MPSORT_ERR_UNSUPPORTED, "key of %lld bytes exceeds 32", (long long)kb);
    cs->elsize = (int64_t)elsize;
    const bool aligned = (((uintptr_t)base | (uintptr_t)out) % (elsize >= 16 ? 16 : elsize)) == 0;
    const bool direct_ok = (elsize == 4 || elsize == 8 || elsize == 16 || elsize == 32) && aligned &&
                           !(flags & MPSORT_GPU_FORCE_GENERAL);
    if (direct_ok) {
        cs->general = false;
        cs->rw = (int)(elsize / 4);
        fill_keyspec(&cs->ks, cs->rw, (int)key->offset, (int)kb, key->width, key->is_signed);
    } else {
        if (kb + 4 > 32) return fail(MPSORT_ERR_UNSUPPORTED, "key of %lld bytes too long for the packed path", (long long)kb);
        cs->general = true;
        cs->rw = kb + 4 <= 8 ? 2 : (kb + 4 <= 16 ? 4 : 8);
        fill_keyspec(&cs->ks, cs->rw, 0, (int)kb, key->width, key->is_signed);
        cs->key_off_user = key->offset;
        cs->key_bytes_user = (int)kb;
    }
    cs->checksum = !(flags & MPSORT_GPU_NO_CHECKSUM);
    cs->timing = (flags & MPSORT_GPU_TIMING) != 0;
    return 0;
}

static void remember_key(DevCtx& c, const mpsort_key_desc* key, int callback, int64_t rsize) {
    c.last_callback = callback;
    if (!callback) {
        c.last_width = key->width;
        c.last_nwords = key->nwords;
        c.last_signed = key->is_signed;
        c.last_rsize = 8 * (int64_t)key->nwords;
    } else {
        c.last_rsize = rsize;
    }
}

// internal key (little-endian words) -> the reference's radix bytes
static void internal_to_radix(const DevCtx& c, const uint64_t* w, unsigned char* radix) {
    unsigned char bytes[32];
    memcpy(bytes, w, 32);
    if (c.last_callback || c.last_width == 8) {
        memcpy(radix, bytes, (size_t)c.last_rsize);
        return;
    }
    for (int q = 0; q < c.last_nwords; q++) {
        uint32_t k;
        memcpy(&k, bytes + 4 * q, 4);
        uint64_t rv;
        if (c.last_signed) {
            const int32_t raw = (int32_t)(k ^ 0x80000000u);
            rv = (uint64_t)(int64_t)raw + 9223372036854775808ULL;
        } else {
            rv = (uint64_t)k;
        }
        memcpy(radix + 8 * q, &rv, 8);
    }
}

}  // namespace mpsg

using namespace mpsg;

extern "C" {

const char* mpsort_gpu_last_error(void) { return g_err; }
const char* mpsort_gpu_version(void) { return "mpsort-mi355x 0.1.0 (gfx950)"; }

int mpsort_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mpsort_gpu_newarray(const void* base, const int64_t* nin, int P, void* out, const int64_t* nout,
                        size_t elsize, const mpsort_key_desc* key, unsigned flags) {
    g_err[0] = 0;
    int64_t ntot = 0;
    int rc = validate_counts(nin, P, nout, &ntot);
    if (rc) return rc;
    CallShape cs;
    memset(&cs, 0, sizeof cs);
    const bool devptr = (flags & MPSORT_GPU_DEVICE_PTRS) != 0;
    rc = shape_for(key, elsize, flags, devptr ? base : nullptr, devptr ? out : nullptr, &cs);
    if (rc) return rc;
    cs.P = P;
    cs.ntot = ntot;
    if (cs.general && ntot >= (int64_t)1 << 32)
        return fail(MPSORT_ERR_UNSUPPORTED, "packed path holds at most 2^32-1 records");
    DevCtx* c = nullptr;
    if ((rc = get_ctx(&c))) return rc;
    mpsort_gpu_stats st;
    memset(&st, 0, sizeof st);
    if (ntot > 0 && (!base || !out)) return fail(MPSORT_ERR_ARG, "NULL data pointer");

    const void* d_in = base;
    void* d_out = out;
    const size_t bytes = (size_t)ntot * elsize;
    hipEvent_t e0 = c->ev[kEvStart], e1 = c->ev[kEvEnd];
    if (!devptr && ntot > 0) {
        if ((rc = c->din.ensure(bytes))) return rc;
        if ((rc = c->dout.ensure(bytes))) return rc;
        MPSG_CHECK(hipEventRecord(e0, c->st));
        MPSG_CHECK(hipMemcpyAsync(c->din.p, base, bytes, hipMemcpyHostToDevice, c->st));
        MPSG_CHECK(hipEventRecord(e1, c->st));
        MPSG_CHECK(hipStreamSynchronize(c->st));
        st.t_h2d_ms = ev_ms(e0, e1);
        d_in = c->din.p;
        d_out = c->dout.p;
    }
    remember_key(*c, key, 0, 0);
    if (ntot > 0) {
        rc = run_histsort(*c, cs, d_in, d_out, nin, nout, nullptr, &st);
    } else {
        st.nvranks = P;
        c->last_P = 0;
    }
    if (!rc && !devptr && ntot > 0) {
        MPSG_CHECK(hipEventRecord(e0, c->st));
        MPSG_CHECK(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, c->st));
        MPSG_CHECK(hipEventRecord(e1, c->st));
        MPSG_CHECK(hipStreamSynchronize(c->st));
        st.t_d2h_ms = ev_ms(e0, e1);
    }
    g_last = st;
    return rc;
}

int mpsort_gpu_newarray_radix(const void* base, const int64_t* nin, int P, void* out, const int64_t* nout,
                              size_t elsize, void (*radix)(const void* ptr, void* radix, void* arg), size_t rsize,
                              void* arg, unsigned flags) {
    g_err[0] = 0;
    int64_t ntot = 0;
    int rc = validate_counts(nin, P, nout, &ntot);
    if (rc) return rc;
    if (!radix) return fail(MPSORT_ERR_ARG, "radix callback is NULL");
    if (rsize == 0 || rsize + 4 > 32) return fail(MPSORT_ERR_UNSUPPORTED, "rsize %zu not in [1, 28]", rsize);
    if (elsize == 0) return fail(MPSORT_ERR_ARG, "elsize is 0");
    if (flags & MPSORT_GPU_DEVICE_PTRS) return fail(MPSORT_ERR_ARG, "the radix-callback form takes host pointers");
    if (ntot >= (int64_t)1 << 32) return fail(MPSORT_ERR_UNSUPPORTED, "at most 2^32-1 records");
    DevCtx* c = nullptr;
    if ((rc = get_ctx(&c))) return rc;
    mpsort_gpu_stats st;
    memset(&st, 0, sizeof st);
    CallShape cs;
    memset(&cs, 0, sizeof cs);
    cs.P = P;
    cs.ntot = ntot;
    cs.general = true;
    cs.elsize = (int64_t)elsize;
    cs.rw = rsize + 4 <= 8 ? 2 : (rsize + 4 <= 16 ? 4 : 8);
    fill_keyspec(&cs.ks, cs.rw, 0, (int)rsize, 0, 0);
    cs.checksum = !(flags & MPSORT_GPU_NO_CHECKSUM);
    cs.timing = (flags & MPSORT_GPU_TIMING) != 0;
    remember_key(*c, nullptr, 1, (int64_t)rsize);
    if (ntot == 0) {
        g_last = st;
        c->last_P = 0;
        return 0;
    }
    // materialise the radix once per record, with the source index
    const size_t rb = (size_t)cs.rw * 4;
    std::vector<uint32_t> packed((size_t)ntot * cs.rw, 0u);
    std::vector<unsigned char> tmp(rsize + 8);
    for (int64_t i = 0; i < ntot; i++) {
        radix((const char*)base + i * elsize, tmp.data(), arg);
        memcpy(&packed[(size_t)i * cs.rw], tmp.data(), rsize);
        packed[(size_t)i * cs.rw + cs.rw - 1] = (uint32_t)i;
    }
    (void)rb;
    const size_t bytes = (size_t)ntot * elsize;
    if ((rc = c->din.ensure(bytes))) return rc;
    if ((rc = c->dout.ensure(bytes))) return rc;
    MPSG_CHECK(hipMemcpy(c->din.p, base, bytes, hipMemcpyHostToDevice));
    // checksum of the input on the host (the device only sees packed records)
    long long sum_in = 0;
    for (size_t b = 0; b < bytes; b++) sum_in += ((const signed char*)base)[b];
    cs.checksum = false;
    rc = run_histsort(*c, cs, c->din.p, c->dout.p, nin, nout, packed.data(), &st);
    if (!rc) {
        MPSG_CHECK(hipMemcpy(out, c->dout.p, bytes, hipMemcpyDeviceToHost));
        long long sum_out = 0;
        for (size_t b = 0; b < bytes; b++) sum_out += ((const signed char*)out)[b];
        st.checksum_in = sum_in;
        st.checksum_out = sum_out;
        if (!(flags & MPSORT_GPU_NO_CHECKSUM) && sum_in != sum_out)
            rc = fail(MPSORT_ERR_CHECKSUM, "MPSort: Data changed after sorting; checksum mismatch");
    }
    g_last = st;
    return rc;
}

int mpsort_gpu_last_stats(mpsort_gpu_stats* s) {
    if (!s) return MPSORT_ERR_ARG;
    *s = g_last;
    return 0;
}

void mpsort_gpu_report_last_run(void) {
    const mpsort_gpu_stats& s = g_last;
    printf("FirstSort: %g\n", s.t_first_sort_ms * 1e-3);
    printf("findP: %g\n", s.t_findp_ms * 1e-3);
    printf("LaySolve: %g\n", s.t_laysolve_ms * 1e-3);
    printf("Exchange: %g\n", s.t_exchange_ms * 1e-3);
    printf("SecondSort: %g\n", s.t_second_sort_ms * 1e-3);
}

void mpsort_mpi_report_last_run() { mpsort_gpu_report_last_run(); }

int mpsort_gpu_last_counts(unsigned char* splitters, unsigned char* pmin, unsigned char* pmax, int64_t* C,
                           int64_t* myCLT, int64_t* myCLE, int64_t* myC, int64_t* send, int64_t* rsize_out) {
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    const int P = c->last_P;
    if (rsize_out) *rsize_out = c->last_rsize;
    if (P <= 0) return fail(MPSORT_ERR_ARG, "no counts recorded (last call sorted nothing)");
    const size_t P1 = (size_t)P + 1;
    MPSG_CHECK(hipStreamSynchronize(c->st));
    if (C) MPSG_CHECK(hipMemcpy(C, c->C, sizeof(int64_t) * P1, hipMemcpyDeviceToHost));
    if (P == 1) {
        // one rank: the matrices are the trivial ones
        if (myCLT) { myCLT[0] = 0; myCLT[1] = c->last_ntot; }
        if (myCLE) { myCLE[0] = 0; myCLE[1] = c->last_ntot; }
        if (myC) { myC[0] = 0; myC[1] = c->last_ntot; }
        if (send) send[0] = c->last_ntot;
    } else {
        if (myCLT) MPSG_CHECK(hipMemcpy(myCLT, c->myclt, sizeof(int64_t) * P * P1, hipMemcpyDeviceToHost));
        if (myCLE) MPSG_CHECK(hipMemcpy(myCLE, c->mycle, sizeof(int64_t) * P * P1, hipMemcpyDeviceToHost));
        if (myC) MPSG_CHECK(hipMemcpy(myC, c->myC, sizeof(int64_t) * P * P1, hipMemcpyDeviceToHost));
        if (send) MPSG_CHECK(hipMemcpy(send, c->send, sizeof(int64_t) * P * P, hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> w((size_t)(P > 1 ? P - 1 : 1) * 4 + 8, 0);
    if (P > 1) {
        MPSG_CHECK(hipMemcpy(w.data(), c->split, sizeof(uint64_t) * 4 * (P - 1), hipMemcpyDeviceToHost));
        if (splitters)
            for (int i = 0; i < P - 1; i++) internal_to_radix(*c, &w[(size_t)i * 4], splitters + (size_t)i * c->last_rsize);
        uint64_t pm[8];
        MPSG_CHECK(hipMemcpy(pm, c->pminmax, sizeof pm, hipMemcpyDeviceToHost));
        if (pmin) internal_to_radix(*c, pm, pmin);
        if (pmax) internal_to_radix(*c, pm + 4, pmax);
    }
    return 0;
}

int mpsort_gpu_set_allocator(mpsort_gpu_malloc_fn m, mpsort_gpu_free_fn f, void* userdata) {
    mpsort_gpu_release();
    g_malloc = m;
    g_free = f;
    g_alloc_ud = userdata;
    return 0;
}

int mpsort_gpu_release(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MPSORT_ERR_NODEVICE;
    DevCtx& c = g_ctx[dev & 63];
    if (!c.init) return 0;
    (void)hipStreamSynchronize(c.st);
    c.din.release();
    c.dout.release();
    c.s0.release();
    c.s1.release();
    c.rbuf.release();
    c.small.release();
    c.lookback.release();
    c.last_P = 0;
    return 0;
}

void mpsort_gpu_set_vranks(int P) { g_vranks = P < 1 ? 1 : P; }
int mpsort_gpu_get_vranks(void) {
    if (g_vranks <= 0) {
        const char* e = getenv("MPSORT_GPU_VRANKS");
        int v = e ? atoi(e) : 1;
        g_vranks = v < 1 ? 1 : v;
    }
    return g_vranks;
}

void* mpsort_gpu_dmalloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}
int mpsort_gpu_dfree(void* p) { return hipFree(p) == hipSuccess ? 0 : MPSORT_ERR_HIP; }
int mpsort_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    MPSG_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int mpsort_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    MPSG_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int mpsort_gpu_dcopy(void* dst, const void* src, size_t bytes) {
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    if (bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0)
        rc = launch_copy(src, dst, (int64_t)(bytes / 16), 4, nullptr, c->st);
    else if (bytes % 4 == 0 && ((uintptr_t)dst % 4) == 0 && ((uintptr_t)src % 4) == 0)
        rc = launch_copy(src, dst, (int64_t)(bytes / 4), 1, nullptr, c->st);
    else
        return fail(MPSORT_ERR_ARG, "dcopy needs 4-byte aligned pointers and sizes");
    if (rc) return rc;
    MPSG_CHECK(hipStreamSynchronize(c->st));
    return 0;
}
int mpsort_gpu_synchronize(void) {
    MPSG_CHECK(hipDeviceSynchronize());
    return 0;
}

// ---- option bits (mpsort-mpi.c:731-767) ----
static void parse_env_options(void) {
    static int parsed = 0;
    if (parsed) return;
    parsed = 1;
    if (getenv("MPSORT_DISABLE_SPARSE_ALLTOALLV")) g_options |= (1 << 1);
    if (getenv("MPSORT_DISABLE_GATHER_SORT")) g_options |= (1 << 3);
    if (getenv("MPSORT_REQUIRE_GATHER_SORT")) g_options |= (1 << 4);
    if (getenv("MPSORT_REQUIRE_SPARSE_ALLTOALLV")) g_options |= (1 << 6);
}
void mpsort_mpi_set_options(int options) {
    parse_env_options();
    g_options |= options;
}
int mpsort_mpi_has_options(int options) {
    parse_env_options();
    return g_options & options;
}
void mpsort_mpi_unset_options(int options) {
    parse_env_options();
    g_options &= ~options;
}

// ---- reference-shaped entry points (mpsort.h:1-11) ----
static void even_split(int64_t n, int P, std::vector<int64_t>& cnt) {
    cnt.resize(P);
    for (int t = 0; t < P; t++) cnt[t] = n * (t + 1) / P - n * t / P;  // mpsort-omp.c:242-243
}

void radix_sort(void* base, size_t nmemb, size_t size, void (*radix)(const void* ptr, void* radix, void* arg),
                size_t rsize, void* arg) {
    const int P = mpsort_gpu_get_vranks();
    std::vector<int64_t> cnt;
    even_split((int64_t)nmemb, P, cnt);
    const int rc = mpsort_gpu_newarray_radix(base, cnt.data(), P, base, cnt.data(), size, radix, rsize, arg, 0);
    if (rc) {
        fprintf(stderr, "radix_sort: %s\n", g_err);
        abort();
    }
}

void mpsort_omp(void* base, size_t nmemb, size_t size, void (*radix)(const void* ptr, void* radix, void* arg),
                size_t rsize, void* arg) {
    radix_sort(base, nmemb, size, radix, rsize, arg);
}

}  // extern "C"
