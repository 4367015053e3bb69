// api.cpp -- C ABI of the MI355X histogram sort: argument checks, device
// workspaces, the histogram-sort pipeline (mpsort-mpi.c:161-604 with P virtual
// ranks), timers, and the reference-compatible entry points of mpsort.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include <algorithm>
#include <map>
#include <string>
#include <chrono>
#include <cmath>

#include "internal.h"
#include "lsort.h"
#include "model.h"
#include "../../include/mpsort_gpu.h"
#include "../../include/mp-mpiu.h"
#include "../../include/mpsort.h"
#include <thread>

namespace mpsg {

// histsort.hip
// model (optional): values tested are the model's interpolated quantiles
// instead of midpoints; exact[i] = 1 when the first one was the splitter
int launch_bisect(const void* recs, int rw, const KeySpec& ks, const int64_t* sbeg,
                  const int64_t* scnt, int P, const int64_t* C, int64_t* win, int64_t* myclt,
                  int64_t* mycle, uint64_t* split, uint64_t* pminmax, int* rounds, uint32_t* err,
                  hipStream_t st, const ModelDev* model = nullptr, int* exact = nullptr, bool perturb = false);
int launch_layout(int P, const int64_t* C, const int64_t* scnt, const int64_t* sbeg, int64_t* myclt,
                  int64_t* mycle, int64_t* myC, int64_t* send, const ExchRuns& ex, uint32_t* err, hipStream_t st);
int launch_exchange(const void* src, void* dst, int64_t ntot, int rw, const RunTable& rt, uint32_t* err,
                    hipStream_t st);
int launch_fault_corrupt(void* byte, long long* checksum_in, hipStream_t st);
int launch_exchange_hist(const void* src, void* dst, int64_t ntot, int rw, const RunTable& rt, uint32_t* err,
                         const int64_t* C, int P, const KeySpec& ks, uint32_t* hist, hipStream_t st);
int launch_pack(const void* in, int64_t n, int64_t elsize, int64_t key_off, int key_bytes,
                void* out, int rw, int iwords, long long* checksum, hipStream_t st);
int launch_gather(const void* in, int64_t n, int64_t elsize, const void* srec, int rw, int iwords, void* out,
                  long long* checksum, uint32_t* err, hipStream_t st);

// Workspace of mpsort_gpu_newarray_noalloc: a fixed array in each GPU's
// memory, held for the life of the process.  Host code never uses this name
// as an address: the device address comes from hipGetSymbolAddress (per
// device, at context set-up) and is kept in DevCtx::scratch.
constexpr size_t kScratchBytes = (size_t)MPSORT_GPU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static char g_err[1024];
static int g_options = 0;
// fault hooks (mpsort_gpu_set_faults): the bits and the bucket capacity
static unsigned g_faults = 0;
static int64_t g_fault_cap = 0;
static uint32_t fault_cap() {
    return (g_faults & MPSORT_GPU_FAULT_BUCKET_CAP) ? (uint32_t)g_fault_cap : (uint32_t)kChunkCap;
}
static int g_vranks = 0;  // 0 = not set yet (environment decides)
static mpsort_gpu_malloc_fn g_malloc = nullptr;
static mpsort_gpu_free_fn g_free = nullptr;
static void* g_alloc_ud = nullptr;
static mpsort_gpu_stats g_last;

// ---- host allocation hook (mp-mpiu.h:4-16): the library's large host
// temporaries go through it
static void* host_malloc_default(const char*, size_t size, const char*, const int, void*) { return malloc(size); }
static void host_free_default(void* p, const char*, const int, void*) { free(p); }
static mpiu_malloc_func g_hmalloc = host_malloc_default;
static mpiu_free_func g_hfree = host_free_default;
static void* g_hud = nullptr;

// host threads for per-record loops over n records (at least 2^16 records
// each): MPSORT_GPU_HOST_THREADS, else the hardware threads up to 16
static int host_threads(int64_t n) {
    static int T = 0;
    if (!T) {
        const char* e = getenv("MPSORT_GPU_HOST_THREADS");
        int v = e ? atoi(e) : (int)std::thread::hardware_concurrency();
        if (!e && v > 16) v = 16;
        T = v < 1 ? 1 : v;
    }
    const int64_t want = (n + (1 << 16) - 1) >> 16;
    return (int)std::max<int64_t>(1, std::min<int64_t>(T, want));
}

// f(t, lo, hi) on T threads over equal ranges of [0, n)
template <class F>
static void parallel_ranges(int64_t n, int T, F f) {
    if (T <= 1) {
        f(0, (int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) th.emplace_back([&f, n, T, t]() { f(t, n / T * t + std::min<int64_t>(t, n % T), n / T * (t + 1) + std::min<int64_t>(t + 1, n % T)); });
    for (auto& x : th) x.join();
}

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
    Buf s2;             // third record buffer (learned-path bucket fallback)
    Buf lsmall;         // learned-path tables and counters
    Buf ovf;            // plans of the learned-path fallback
    Buf small;          // counts, tables
    Buf lookback;       // look-back flags (only ever holds flag words)
    Buf ltile;          // learned scatter: per-tile bucket counts and offsets
    Buf mbnd;           // merged SecondSort: sub-run bounds per (receiver, sender, bucket)
    Buf mdesc;          // merged SecondSort: group descriptors and their sub-run tables
    Buf fsamp;          // plain bucket FirstSort: the segments' sample keys (three buffers)
    Buf fsplan;         // plain bucket FirstSort: the sample sort's plan arrays
    int64_t fsplan_key = 0;  // segments and tile shape those arrays were made for (0: none)
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    Buf hmirror;        // device copy of staged host tables (Upload)
    Buf gstage;         // small device results gathered for one download (Readback)
    void* rpin = nullptr;  // pinned: per-call staging and small reads (pin_take), reset per call
    size_t rpin_bytes = 0, rpin_off = 0;
    void* scratch = nullptr;      // this device's copy of the scratch array
    hipError_t scratch_err = hipSuccess;
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
        // the phase and pass events only time the work (the host waits on
        // the stream): no system-scope fence, which would write back and
        // invalidate the caches between kernels at every event
        for (int i = 0; i < kEvPhase; i++) MPSG_CHECK(hipEventCreateWithFlags(&c.ev[i], hipEventDisableSystemFence));
        for (int a = 0; a < 2; a++)
            for (int i = 0; i < 2 * kMaxPasses; i++)
                MPSG_CHECK(hipEventCreateWithFlags(&c.pass_ev[a][i], hipEventDisableSystemFence));
        c.lookback.zero_on_grow = true;
        // the device address of this GPU's scratch array (the device is
        // current); only the allocation-free entry point uses it
        c.scratch_err = hipGetSymbolAddress(&c.scratch, HIP_SYMBOL(scratch));
        if (c.scratch_err != hipSuccess) c.scratch = nullptr;
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

// the per-call pinned area: staging for host tables and small reads, taken
// in order and reset at the start of every call (the previous call's copies
// have all completed by then)
static int ensure_rpin(DevCtx& c, size_t bytes) {
    c.rpin_off = 0;
    if (bytes <= c.rpin_bytes) return 0;
    if (c.rpin) (void)hipHostFree(c.rpin);
    c.rpin = nullptr;
    c.rpin_bytes = 0;
    const size_t want = bytes + 4096;
    MPSG_CHECK(hipHostMalloc(&c.rpin, want, hipHostMallocDefault));
    c.rpin_bytes = want;
    return 0;
}
static void* pin_take(DevCtx& c, size_t bytes) {
    const size_t at = (c.rpin_off + 63) & ~(size_t)63;
    if (!c.rpin || at + bytes > c.rpin_bytes) return nullptr;
    c.rpin_off = at + bytes;
    return (char*)c.rpin + at;
}

// Host tables bound for device arrays and small counters to clear: the
// tables lie in one pinned buffer (base); flush() copies their span to the
// device mirror once and one fan-out launch writes every array and fill,
// instead of a copy or a launch per array.  The host tables must stay as
// they are until the stream has run the copy.
struct Upload {
    struct Item {
        uint32_t* dst;
        size_t at;  // byte offset in base, or ~0 for a fill
        uint32_t nwords, fill;
    };
    DevCtx& c;
    hipStream_t st;
    const char* base;
    size_t base_bytes;
    Item it[kFanMax];
    int n = 0;
    size_t lo = ~(size_t)0, hi = 0;
    Upload(DevCtx& c_, hipStream_t st_, const void* b, size_t bb)
        : c(c_), st(st_), base((const char*)b), base_bytes(bb) {}
    int add(void* dst, const void* hsrc, size_t bytes) {
        const size_t at = (size_t)((const char*)hsrc - base);
        if (!base || (const char*)hsrc < base || at + bytes > base_bytes || at % 4 || bytes % 4)
            return fail(MPSORT_ERR_ARG, "host table outside the staging buffer");
        if (n == kFanMax) {
            int rc = flush();
            if (rc) return rc;
        }
        it[n++] = Item{(uint32_t*)dst, at, (uint32_t)(bytes / 4), 0u};
        lo = std::min(lo, at);
        hi = std::max(hi, at + bytes);
        return 0;
    }
    int fill(uint32_t* dst, uint32_t v, int64_t nw) {
        if (nw <= 0) return 0;
        if (nw > (1 << 20)) return launch_fill_u32(dst, v, nw, st);
        if (n == kFanMax) {
            int rc = flush();
            if (rc) return rc;
        }
        it[n++] = Item{dst, ~(size_t)0, (uint32_t)nw, v};
        return 0;
    }
    int flush() {
        int rc = 0;
        if (hi > lo) {
            if ((rc = c.hmirror.ensure(base_bytes))) return rc;
            MPSG_CHECK(hipMemcpyAsync((char*)c.hmirror.p + lo, base + lo, hi - lo, hipMemcpyHostToDevice, st));
        }
        FanOut f;
        f.n = n;
        for (int i = 0; i < n; i++)
            f.e[i] = FanEntry{it[i].dst,
                              it[i].at == ~(size_t)0 ? nullptr : (const uint32_t*)((const char*)c.hmirror.p + it[i].at),
                              it[i].nwords, it[i].fill};
        if (n && (rc = launch_fanout(f, st))) return rc;
        n = 0;
        lo = ~(size_t)0;
        hi = 0;
        return 0;
    }
};

// Small device-to-host reads: with use, wait() gathers them into one device
// area with one fan-out launch, copies that area into the per-call pinned
// area once and waits once; otherwise (or when they do not fit) each copy
// goes straight to its target (a copy into pageable memory waits by itself).
struct Readback {
    DevCtx& c;
    hipStream_t st;
    bool use;
    struct Item {
        void* host;
        const void* dev;
        size_t bytes;
    };
    Item it[kFanMax];
    int n = 0;
    Readback(DevCtx& c_, hipStream_t st_, bool use_) : c(c_), st(st_), use(use_) {}
    int add(void* host, const void* dev, size_t bytes) {
        if (!use || n == kFanMax || bytes % 4 || (uintptr_t)dev % 4) {
            MPSG_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
            return 0;
        }
        it[n++] = Item{host, dev, bytes};
        return 0;
    }
    int wait() {
        int rc = 0;
        size_t total = 0;
        for (int i = 0; i < n; i++) total += it[i].bytes;
        char* pin = n ? (char*)pin_take(c, total) : nullptr;
        if (n && pin && !(rc = c.gstage.ensure(total))) {
            FanOut f;
            f.n = n;
            size_t off = 0;
            for (int i = 0; i < n; i++) {
                f.e[i] = FanEntry{(uint32_t*)((char*)c.gstage.p + off), (const uint32_t*)it[i].dev,
                                  (uint32_t)(it[i].bytes / 4), 0u};
                off += it[i].bytes;
            }
            if ((rc = launch_fanout(f, st))) return rc;
            MPSG_CHECK(hipMemcpyAsync(pin, c.gstage.p, total, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipStreamSynchronize(st));
            off = 0;
            for (int i = 0; i < n; i++) {
                memcpy(it[i].host, pin + off, it[i].bytes);
                off += it[i].bytes;
            }
        } else {
            if (rc) return rc;
            for (int i = 0; i < n; i++)
                MPSG_CHECK(hipMemcpyAsync(it[i].host, it[i].dev, it[i].bytes, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipStreamSynchronize(st));
        }
        n = 0;
        return 0;
    }
};

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
    PlanArrays unused;  // an Arena given no target still takes the space
    if (!pa) pa = &unused;
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

// the compact run table of the sparse exchange: at most this many runs per
// receiver on average (more: the dense table is used)
constexpr int kSparseRunsPerRank = 64;
static size_t sparse_cap(int P) { return std::min<size_t>((size_t)P * P, (size_t)P * kSparseRunsPerRank); }

struct HistArrays {
    int64_t *C, *win, *myclt, *mycle, *myC, *send, *run_dst, *run_src, *run_dst_c, *run_src_c;
    int *rnz, *snz, *roff, *sel;
    uint64_t *split, *pminmax;
    int* rounds;
    uint32_t* err;
    long long* checksum;
};

template <class A>
static void lay_hist(A& a, int P, HistArrays* h) {
    HistArrays unused;  // an Arena given no target still takes the space
    if (!h) h = &unused;
    const size_t P1 = (size_t)P + 1, PP = (size_t)P * P, Pm = P > 1 ? (size_t)P - 1 : 1;
    if constexpr (std::is_same<A, Arena>::value) {
        h->C = a.template take<int64_t>(P1);
        h->win = a.template take<int64_t>(Pm * P * 6);
        h->myclt = a.template take<int64_t>(P * P1);
        h->mycle = a.template take<int64_t>(P * P1);
        h->myC = a.template take<int64_t>(P * P1);
        h->send = a.template take<int64_t>(PP);
        h->run_dst = a.template take<int64_t>(PP);
        h->run_src = a.template take<int64_t>(PP);
        h->run_dst_c = a.template take<int64_t>(sparse_cap(P));
        h->run_src_c = a.template take<int64_t>(sparse_cap(P));
        h->rnz = a.template take<int>(P);
        h->snz = a.template take<int>(P);
        h->roff = a.template take<int>(P1);
        h->sel = a.template take<int>(4);
        h->split = a.template take<uint64_t>(Pm * 4);
        h->pminmax = a.template take<uint64_t>(8);
        h->rounds = a.template take<int>(Pm);
        h->err = a.template take<uint32_t>(4);
        h->checksum = a.template take<long long>(2);
    } else {
        a.template take<int64_t>(P1);
        a.template take<int64_t>(Pm * P * 6);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(P * P1);
        a.template take<int64_t>(PP);
        a.template take<int64_t>(PP);
        a.template take<int64_t>(PP);
        a.template take<int64_t>(sparse_cap(P));
        a.template take<int64_t>(sparse_cap(P));
        a.template take<int>(P);
        a.template take<int>(P);
        a.template take<int>(P1);
        a.template take<int>(4);
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
    int iwords;            // general path: words of the source index (1: u32, 2: u64)
    bool checksum;
    bool timing;
    bool exchange;  // keep the explicit exchange before the SecondSort
    int exch_policy;  // ExchPolicy of the exchange's run table
    int first_mode;   // plain FirstSort: 0 by size, 1 radix passes, 2 buckets
};

static ExchRuns exch_runs(const HistArrays& h, int P, int policy) {
    ExchRuns e;
    e.dst = h.run_dst;
    e.src = h.run_src;
    e.dst_c = h.run_dst_c;
    e.src_c = h.run_src_c;
    e.rnz = h.rnz;
    e.snz = h.snz;
    e.roff = h.roff;
    e.sel = h.sel;
    e.cap = (int)sparse_cap(P);
    e.policy = policy;
    return e;
}
static RunTable run_table(const HistArrays& h, int P) {
    return RunTable{h.run_dst, h.run_src, h.run_dst_c, h.run_src_c, h.sel, P * P};
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

constexpr int kMaxOverflowList = 1 << 16;  // bucket overflow list capacity
constexpr int kMaxFallbackSegs = 4096;     // larger lists re-sort the whole side plainly

// Device workspaces of one plain histogram sort and their sizes.  The
// sort-record buffers hold n sort records each; s0 is needed with more than
// one rank or packed records, rbuf with packed records only.
struct HistNeed {
    size_t small = 0, s0 = 0, s1 = 0, rbuf = 0, lookback = 0;
};
struct HistWork {
    char* small = nullptr;
    void *s0 = nullptr, *s1 = nullptr, *rbuf = nullptr;
    uint64_t* lookback = nullptr;
    int64_t* host_tab = nullptr;   // 7 P + 4 int64 (tables staged to the device)
    bool clear_lookback = false;   // the flag words may hold other data: clear them
    bool merge = false;            // the merged SecondSort may allocate its tables
};

static int64_t tiles_of(const int64_t* counts, int P, int64_t T) {
    int64_t t = 0;
    for (int i = 0; i < P; i++) t += (counts[i] + T - 1) / T;
    return t;
}

// merge: the flag words also serve the merged SecondSort's bucket fallback
static HistNeed histsort_need(const CallShape& cs, const int64_t* nin, const int64_t* nout, bool merge = false) {
    HistNeed w;
    Sizer sz;
    lay_plan(sz, cs.P, cs.ks.key_bytes, nullptr);
    lay_plan(sz, cs.P, cs.ks.key_bytes, nullptr);
    lay_hist(sz, cs.P, nullptr);
    w.small = sz.off + 4096;
    const size_t rec = (size_t)cs.ntot * (size_t)cs.rw * 4;
    w.s0 = (cs.P > 1 || cs.general) ? rec : 0;
    w.s1 = rec;
    w.rbuf = cs.general ? rec : 0;
    const TileCfg tile = default_tile(cs.rw);
    const int64_t T = (int64_t)tile.block * tile.items;
    const int64_t nt = std::max<int64_t>(1, std::max(tiles_of(nin, cs.P, T), tiles_of(nout, cs.P, T)));
    w.lookback = (size_t)(nt + (merge ? kMaxFallbackSegs + 1 : 0)) * kBins * sizeof(uint64_t);
    return w;
}

// the per-device cache, grown to the need
static int histwork_cached(DevCtx& c, const CallShape& cs, const HistNeed& need, HistWork* w) {
    int rc = 0;
    if ((rc = c.small.ensure(need.small))) return rc;
    if (need.s0 && (rc = c.s0.ensure(need.s0))) return rc;
    if ((rc = c.s1.ensure(need.s1))) return rc;
    if (need.rbuf && (rc = c.rbuf.ensure(need.rbuf))) return rc;
    if ((rc = c.lookback.ensure(need.lookback))) return rc;
    if ((rc = ensure_pinned(c, (size_t)(7 * cs.P + 4) * sizeof(int64_t) + 4096))) return rc;
    w->small = (char*)c.small.p;
    w->s0 = c.s0.p;
    w->s1 = c.s1.p;
    w->rbuf = c.rbuf.p;
    w->lookback = (uint64_t*)c.lookback.p;
    w->host_tab = (int64_t*)c.pinned;
    w->clear_lookback = false;
    w->merge = true;
    return 0;
}

// The SecondSort gathers every receiver bucket from the senders' sorted
// segments, with no exchange, when no exchange is asked for: at most
// kMergeMaxPH ranks, records of 8 or 16 bytes with keys of at most 8 bytes,
// positions inside a segment below 2^31.
static bool merge_eligible(const CallShape& cs) {
    return !cs.exchange && cs.P > 1 && cs.P <= kMergeMaxPH && (cs.rw == 2 || cs.rw == 4) && cs.ks.key_bytes <= 8 &&
           cs.ntot < ((int64_t)1 << 31);
}
// The plain FirstSort is a bucket sort (tables from a sample of every
// segment) for segments of keys up to 8 bytes in 8- or 16-byte sort records,
// each at most half of 1024 in-LDS buckets, at most kFirstMaxP of them; by
// default only from kFirstMinAvg records per segment on.
constexpr int kFirstMaxP = 256;
constexpr int64_t kFirstMinAvg = (int64_t)1 << 18;
constexpr int kFirstSamples = 65536;  // sample keys per segment (64 per bucket)
static bool first_buckets_eligible(const CallShape& cs, const int64_t* nin) {
    if (cs.first_mode == 1 || cs.P > kFirstMaxP || !(cs.rw == 2 || cs.rw == 4) || cs.ks.key_bytes > 8 ||
        cs.ntot >= ((int64_t)1 << 31))
        return false;
    int64_t big = 0;
    for (int t = 0; t < cs.P; t++) big = std::max(big, nin[t]);
    if (big > (int64_t)kLBins * kChunkCap / 2) return false;
    return cs.first_mode == 2 || cs.ntot >= kFirstMinAvg * cs.P;
}
// what the plain path's merged SecondSort (and bucket FirstSort) reports
struct MergedReport {
    uint32_t lists[3] = {0, 0, 0};  // oversized, LDS radix, 1024-thread form (read at the call's end)
    const uint32_t* lists_dev = nullptr;  // where they are on the device
    const uint32_t* flag_dev = nullptr;   // a device error flag to check at the call's end, or null
    int overflow = 0;
    int global_fallback = 0;
};
static int plain_merged_second(DevCtx& c, const CallShape& cs, const SegSortPlan& planB, const HistArrays& h,
                               const PlanArrays& pa, const void* src, void* dst, void* tmp, long long* cs_out,
                               hipEvent_t* ev, MergedReport* rep, hipStream_t st);
static int plain_bucket_first(DevCtx& c, const CallShape& cs, const SegSortPlan& planA, const int64_t* hbegin,
                              const int64_t* hcount, const void* src, void* dst, void* tmp, long long* cs_in,
                              long long* cs_out, hipEvent_t* ev, MergedReport* rep, hipStream_t st);

// The histogram sort on device-resident data.  d_in / d_out are the caller's
// records (device); for the general path they are packed into sort records.
static int run_histsort(DevCtx& c, const CallShape& cs, const void* d_in, void* d_out,
                        const int64_t* nin, const int64_t* nout, const void* packed_host,
                        mpsort_gpu_stats* stats, const HistWork& w) {
    hipStream_t st = c.st;
    const int P = cs.P;
    const int rw = cs.rw;
    const int64_t n = cs.ntot;
    const size_t recbytes = (size_t)rw * 4;
    const int npass = cs.ks.key_bytes;
    int rc = 0;

    // ---- workspace
    Arena ar{w.small};
    PlanArrays pa, pb;
    HistArrays h;
    lay_plan(ar, P, npass, &pa);
    lay_plan(ar, P, npass, &pb);
    lay_hist(ar, P, &h);

    // host tables: [beginA, countA, tsA(P+1), beginB, countB, tsB(P+1), C(P+1)]
    int64_t* ht = w.host_tab;
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
    if (w.clear_lookback &&
        (rc = launch_fill_u32((uint32_t*)w.lookback, 0u, (int64_t)(ntiles > 0 ? ntiles : 1) * kBins * 2, st)))
        return rc;

    if (w.merge) {
        // the allocating entry points: the tables are in the pinned buffer,
        // uploaded by one copy and one fan-out launch with the clears
        if ((rc = ensure_rpin(c, (size_t)64 * (P + 16) * sizeof(int64_t)))) return rc;
        Upload up(c, st, c.pinned, c.pinned_bytes);
        if ((rc = up.add(pa.begin, hbA, sizeof(int64_t) * P)) || (rc = up.add(pa.count, hcA, sizeof(int64_t) * P)) ||
            (rc = up.add(pa.tile_start, htA, sizeof(int64_t) * (P + 1))) ||
            (rc = up.add(pb.begin, hbB, sizeof(int64_t) * P)) || (rc = up.add(pb.count, hcB, sizeof(int64_t) * P)) ||
            (rc = up.add(pb.tile_start, htB, sizeof(int64_t) * (P + 1))) ||
            (rc = up.add(h.C, hC, sizeof(int64_t) * (P + 1))) || (rc = up.fill(h.err, 0u, 4)) ||
            (rc = up.fill((uint32_t*)h.checksum, 0u, 4)) || (rc = up.fill(pa.pass_flags, 0u, kMaxPasses + 1)) ||
            (rc = up.fill(pb.pass_flags, 0u, kMaxPasses + 1)) || (rc = up.flush()))
            return rc;
    } else {
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
    }

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
        p.d_lookback = w.lookback;
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
            MPSG_CHECK(hipMemcpyAsync(w.rbuf, packed_host, n * recbytes, hipMemcpyHostToDevice, st));
        } else {
            if ((rc = launch_pack(d_in, n, cs.elsize, cs.key_off_user, cs.key_bytes_user, w.rbuf, rw, cs.iwords,
                                  cs.checksum ? h.checksum : nullptr, st)))
                return rc;
        }
        recs_in = w.rbuf;
        final_dst = w.rbuf;
    }
    const bool cs_direct = cs.checksum && !cs.general;
    // the corruption fault hook changes a byte past the key of the first
    // FirstSort output record (more than one rank, direct records), else
    // the recorded input sum
    const bool corrupt_record = P > 1 && !cs.general && cs.ks.key_off + cs.ks.key_bytes < rw * 4;

    int first_passes = 0, second_passes = 0;
    hipEvent_t* pevA = cs.timing ? c.pass_ev[0] : nullptr;
    hipEvent_t* pevB = cs.timing ? c.pass_ev[1] : nullptr;
    const bool merged = w.merge && merge_eligible(cs);
    const bool first_buckets = w.merge && first_buckets_eligible(cs, hcA);
    MergedReport mrep, frep;
    if (P == 1) {
        // one rank: FirstSort is the whole sort (the exchange is the identity
        // and SecondSort re-sorts a sorted run)
        if (first_buckets) {
            if ((rc = plain_bucket_first(c, cs, planA, hbA, hcA, recs_in, final_dst, w.s1,
                                         cs_direct ? h.checksum : nullptr, cs_direct ? h.checksum + 1 : nullptr, pevA,
                                         &frep, st)))
                return rc;
        } else {
            SortBuffers b{recs_in, final_dst, w.s1};
            if ((rc = segmented_sort(planA, b, st, cs_direct, cs_direct, &first_passes, pevA))) return rc;
        }
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        for (int e = kEvFindP; e <= kEvSecond; e++) MPSG_CHECK(hipEventRecord(c.ev[e], st));
    } else {
        if (first_buckets) {
            if ((rc = plain_bucket_first(c, cs, planA, hbA, hcA, recs_in, w.s0, w.s1,
                                         cs_direct ? h.checksum : nullptr, nullptr, pevA, &frep, st)))
                return rc;
        } else {
            SortBuffers b1{recs_in, w.s0, w.s1};
            if ((rc = segmented_sort(planA, b1, st, cs_direct, false, &first_passes, pevA))) return rc;
        }
        if ((g_faults & MPSORT_GPU_FAULT_CORRUPT) && corrupt_record &&
            (rc = launch_fault_corrupt((char*)w.s0 + rw * 4 - 1, nullptr, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        if ((rc = launch_bisect(w.s0, rw, cs.ks, pa.begin, pa.count, P, h.C, h.win, h.myclt, h.mycle,
                                h.split, h.pminmax, h.rounds, h.err, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFindP], st));
        if ((rc = launch_layout(P, h.C, pa.count, pa.begin, h.myclt, h.mycle, h.myC, h.send,
                                exch_runs(h, P, cs.exch_policy), h.err, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvLayout], st));
        if (merged) {
            // every receiver bucket gathered from the senders' sorted
            // segments and sorted in LDS (receiver tables from a sample)
            if ((rc = plain_merged_second(c, cs, planB, h, pa, w.s0, final_dst, w.s1,
                                          cs_direct ? h.checksum + 1 : nullptr, pevB, &mrep, st)))
                return rc;
        } else {
            // the exchange also builds the SecondSort's digit histograms
            if ((rc = seg_clear(planB, st))) return rc;
            if ((rc = launch_exchange_hist(w.s0, w.s1, n, rw, run_table(h, P), h.err, h.C, P, cs.ks,
                                           planB.d_hist, st)))
                return rc;
            MPSG_CHECK(hipEventRecord(c.ev[kEvExchange], st));
            SortBuffers b2{w.s1, final_dst, w.s0};
            if ((rc = seg_passes(planB, b2, st, cs_direct ? h.checksum + 1 : nullptr, &second_passes, pevB)))
                return rc;
        }
        MPSG_CHECK(hipEventRecord(c.ev[kEvSecond], st));
    }
    if (cs.general) {
        if ((rc = launch_gather(d_in, n, cs.elsize, w.rbuf, rw, cs.iwords, d_out,
                                cs.checksum ? h.checksum + 1 : nullptr, h.err, st)))
            return rc;
    }
    if ((g_faults & MPSORT_GPU_FAULT_CORRUPT) && !corrupt_record &&
        (rc = launch_fault_corrupt(nullptr, h.checksum, st)))
        return rc;
    MPSG_CHECK(hipEventRecord(c.ev[kEvEnd], st));

    // ---- results
    uint32_t errs[3] = {0, 0, 0};
    long long sums[2] = {0, 0};
    int rounds_host = 0;
    static int rounds[4096];
    int sel[2] = {0, 0};
    uint32_t sample_flag = 0;
    {
        Readback rb(c, st, w.merge);
        if ((rc = rb.add(&errs[0], h.err, sizeof(uint32_t))) ||
            (rc = rb.add(&errs[1], pa.pass_flags + kMaxPasses, sizeof(uint32_t))) ||
            (rc = rb.add(&errs[2], pb.pass_flags + kMaxPasses, sizeof(uint32_t))) ||
            (rc = rb.add(sums, h.checksum, sizeof sums)))
            return rc;
        if (P > 1 && ((rc = rb.add(rounds, h.rounds, sizeof(int) * (P - 1))) || (rc = rb.add(sel, h.sel, sizeof sel))))
            return rc;
        for (MergedReport* r : {&mrep, &frep})
            if (r->lists_dev && (rc = rb.add(r->lists, r->lists_dev, sizeof r->lists))) return rc;
        if (frep.flag_dev && (rc = rb.add(&sample_flag, frep.flag_dev, sizeof sample_flag))) return rc;
        if ((rc = rb.wait())) return rc;
    }
    if (sample_flag) return fail(MPSORT_ERR_TIMEOUT, "FirstSort sample sort: device flag 0x%x", sample_flag);
    for (int i = 0; i + 1 < P; i++) rounds_host = rounds[i] > rounds_host ? rounds[i] : rounds_host;

    // stats
    stats->nvranks = P;
    stats->exchange_sparse = merged ? 0 : sel[0];
    stats->exchange_runs = P > 1 && !merged ? (sel[0] ? sel[1] : P * P) : 0;
    stats->second_merged = P > 1 && merged ? 1 : 0;
    if (P > 1 && merged) {
        stats->buckets = P * kMergeBins;
        stats->buckets_overflow = mrep.overflow;
        stats->buckets_lds_radix = (int32_t)mrep.lists[1];
        stats->buckets_listed = (int32_t)mrep.lists[2];
        stats->plain_sides = mrep.global_fallback;
    }
    if (first_buckets) {
        stats->buckets += P * kLBins;
        stats->buckets_overflow += frep.overflow;
        stats->buckets_overflow_first = frep.overflow;
        stats->buckets_lds_radix += (int32_t)frep.lists[1];
        stats->buckets_listed += (int32_t)frep.lists[2];
        stats->plain_sides += frep.global_fallback;
    }
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
        if (first_buckets) {
            // sample and tables; bucket histogram, scatter, in-LDS bucket
            // sort, LDS radix sort of crowded buckets, overflow fallback
            hipEvent_t* e = c.pass_ev[0];
            stats->pass_ms[np++] = ev_ms(e[8], e[0]);
            for (int k = 0; k < 5; k++) stats->pass_ms[np++] = ev_ms(e[k], e[k + 1]);
        }
        for (int i = 0; i < second_passes && np < 64; i++)
            stats->pass_ms[np++] = ev_ms(c.pass_ev[1][2 * i], c.pass_ev[1][2 * i + 1]);
        if (P > 1 && merged) {
            // tables, bounds, counts and groups; the gather-and-sort launches;
            // the listed buckets; the overflow fallback
            hipEvent_t* e = c.pass_ev[1];
            stats->pass_ms[np++] = ev_ms(e[8], e[6]);
            stats->pass_ms[np++] = ev_ms(e[6], e[7]);
            stats->pass_ms[np++] = ev_ms(e[3], e[4]);
            stats->pass_ms[np++] = ev_ms(e[4], e[5]);
        }
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
    // with host-built sort records the caller compares the host's input sum
    if (cs.checksum && !packed_host && sums[0] != sums[1])
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

// packed sort records of key_bytes key bytes and the source index: words
// (2, 4 or 8), or 0 when they do not fit 32 bytes
static int packed_words(int64_t key_bytes, int iwords) {
    const int64_t b = key_bytes + 4 * iwords;
    return b <= 8 ? 2 : (b <= 16 ? 4 : (b <= 32 ? 8 : 0));
}

static int shape_for(const mpsort_key_desc* key, size_t elsize, unsigned flags, const void* base, const void* out,
                     int64_t ntot, CallShape* cs) {
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
        cs->iwords = ntot >= ((int64_t)1 << 32) ? 2 : 1;
        cs->rw = packed_words(kb, cs->iwords);
        if (!cs->rw)
            return fail(MPSORT_ERR_UNSUPPORTED, "key of %lld bytes and a %d-byte index exceed the 32-byte packed record",
                        (long long)kb, 4 * cs->iwords);
        cs->general = true;
        fill_keyspec(&cs->ks, cs->rw, 0, (int)kb, key->width, key->is_signed);
        cs->key_off_user = key->offset;
        cs->key_bytes_user = (int)kb;
    }
    cs->checksum = !(flags & MPSORT_GPU_NO_CHECKSUM);
    cs->timing = (flags & MPSORT_GPU_TIMING) != 0;
    cs->exchange = (flags & MPSORT_GPU_LEARNED_EXCHANGE) != 0;
    // the call's flags, else the recorded options (REQUIRE over DISABLE, as
    // mpsort-mpi.c:576-583), else the vote
    if (flags & MPSORT_GPU_EXCHANGE_SPARSE) cs->exch_policy = kExchSparse;
    else if (flags & MPSORT_GPU_EXCHANGE_DENSE) cs->exch_policy = kExchDense;
    else if (g_options & MPSORT_REQUIRE_SPARSE_ALLTOALLV) cs->exch_policy = kExchSparse;
    else if (g_options & MPSORT_DISABLE_SPARSE_ALLTOALLV) cs->exch_policy = kExchDense;
    else cs->exch_policy = kExchAuto;
    // a run table chosen by the caller means an exchange
    if (cs->exch_policy != kExchAuto) cs->exchange = true;
    cs->first_mode = (flags & MPSORT_GPU_FIRST_RADIX) ? 1 : ((flags & MPSORT_GPU_FIRST_BUCKETS) ? 2 : 0);
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


// ===========================================================================
// learned-splitter path

struct LearnedModel {
    int device = 0;
    float* raw = nullptr;     // kMlpParams
    float* eff = nullptr;     // kMlpParams
    float* leaves = nullptr;  // 4 * L
    float* xstart = nullptr;  // L + 1
    uint64_t* heavy = nullptr;  // duplicate-heavy keys, ascending
    float2* heavy_q = nullptr;  // their quantile ranges in the training sample
    int nheavy = 0;
    int L = 0;
    int linear_root = 0;
    uint64_t kmin = 0, kmax = 0;
    double inv_range = 0;
    int log_input = 0;
    double log_scale = 0;
    int64_t nsample = 0;
    float err_max = 0, loss = 0;
    double train_ms = 0;
    int steps = 0;
    int max_leaf = 0;
    int key_bytes = 8;
    ModelDev dev() const {
        ModelDev m;
        m.eff = eff;
        m.leaves = leaves;
        m.xstart = xstart;
        m.kmax = kmax;
        m.L = L;
        m.linear_root = linear_root;
        m.kmin = kmin;
        m.inv_range = inv_range;
        m.log_input = log_input;
        m.log_scale = log_scale;
        return m;
    }
};


struct LearnedSide {
    int64_t *begin, *count, *ltiles;  // segment tables (learned tile size)
    uint32_t* lhist;                  // nseg * buckets per segment
    int64_t* bbase;                   // nseg * buckets per segment
    uint64_t* tables;                 // nseg * boundaries per segment
    float *q0, *q1;                   // nseg
    int* nedge;                       // nseg: edge boundaries of each table (0: none)
    uint32_t* tickets;                // 4
    uint32_t* err;                    // 4
    int64_t* overflow;                // 2 * kMaxOverflowList
    uint32_t* noverflow;              // 8: [0] oversized buckets, [1] buckets for the LDS radix sort,
                                      //    [2] medium, [3] groups, [4] larger single buckets
                                      //    (merged SecondSort)
    uint32_t* hard;                   // nseg * buckets per segment
    uint32_t* medium;                 // nseg * buckets per segment (merged SecondSort)
    uint32_t* tasks;                  // 2 * nseg * buckets per segment (merged SecondSort groups, then
                                      // larger single buckets)
    int64_t tstride = kLBounds;       // segment s uses tables + s * tstride (0: one shared table)
};

// sized for the larger of the two bucket counts (kLBins, kMergeBins per segment)
template <class A>
static void lay_lside(A& a, int P, LearnedSide* ls) {
    LearnedSide unused;  // an Arena given no target still takes the space
    if (!ls) ls = &unused;
    if constexpr (std::is_same<A, Arena>::value) {
        ls->begin = a.template take<int64_t>(P);
        ls->count = a.template take<int64_t>(P);
        ls->ltiles = a.template take<int64_t>(P + 1);
        ls->lhist = a.template take<uint32_t>((size_t)P * kMergeBins);
        ls->bbase = a.template take<int64_t>((size_t)P * kMergeBins);
        ls->tables = a.template take<uint64_t>((size_t)P * kMergeBounds);
        ls->q0 = a.template take<float>(P);
        ls->q1 = a.template take<float>(P);
        ls->nedge = a.template take<int>(P);
        ls->tickets = a.template take<uint32_t>(4);
        ls->err = a.template take<uint32_t>(4);
        ls->overflow = a.template take<int64_t>(2 * (size_t)kMaxOverflowList);
        ls->noverflow = a.template take<uint32_t>(8);
        ls->hard = a.template take<uint32_t>((size_t)P * kMergeBins);
        ls->medium = a.template take<uint32_t>((size_t)P * kMergeBins);
        ls->tasks = a.template take<uint32_t>(2 * (size_t)P * kMergeBins);
    } else {
        a.template take<int64_t>(P);
        a.template take<int64_t>(P);
        a.template take<int64_t>(P + 1);
        a.template take<uint32_t>((size_t)P * kMergeBins);
        a.template take<int64_t>((size_t)P * kMergeBins);
        a.template take<uint64_t>((size_t)P * kMergeBounds);
        a.template take<float>(P);
        a.template take<float>(P);
        a.template take<int>(P);
        a.template take<uint32_t>(4);
        a.template take<uint32_t>(4);
        a.template take<int64_t>(2 * (size_t)kMaxOverflowList);
        a.template take<uint32_t>(8);
        a.template take<uint32_t>((size_t)P * kMergeBins);
        a.template take<uint32_t>((size_t)P * kMergeBins);
        a.template take<uint32_t>(2 * (size_t)P * kMergeBins);
    }
}

struct LearnedCounters {
    const uint32_t* lists[2] = {nullptr, nullptr};  // device list counters of the sides that ran
    int overflow_buckets = 0;
    int global_fallback = 0;
    int unbucketed = 0;  // sides sorted without the bucket histogram (no bucket bases)
};

// Radix sort of the listed segments (device list of nov {begin, count}
// pairs) from src into dst at the same positions; tmp is the third buffer
// of the digit passes.  plain supplies the record shape and key.  Segments
// of one key value are in stable order already: copied, not sorted.
static int sort_listed_segments(DevCtx& c, const SegSortPlan& plain, int rw, int64_t* d_list, uint32_t nov,
                                const void* src, void* dst, void* tmp, long long* cs_out, hipStream_t st) {
    int rc = 0, passes = 0;
    std::vector<int64_t> ov(2 * (size_t)nov);
    MPSG_CHECK(hipMemcpyAsync(ov.data(), d_list, sizeof(int64_t) * 2 * nov, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    int64_t maxc = 0;
    for (uint32_t i = 0; i < nov; i++) maxc = std::max(maxc, ov[2 * i + 1]);
    if ((rc = launch_one_value_filter(rw, src, dst, make_lkey(plain.ks), d_list, nov, maxc, cs_out, st))) return rc;
    MPSG_CHECK(hipMemcpyAsync(ov.data(), d_list, sizeof(int64_t) * 2 * nov, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    // the segments left to sort: those flagged (top bit) as holding several
    // key values
    std::vector<int> ord;
    for (int i = 0; i < (int)nov; i++)
        if (ov[2 * i + 1] < 0) {
            ov[2 * i + 1] &= INT64_MAX;
            ord.push_back(i);
        }
    const int ns = (int)ord.size();
    if (ns == 0) return 0;
    std::vector<int64_t> ob(ns), oc(ns), ot(ns + 1), ob2(ns);
    // segments must be in ascending position order for the tile tables
    std::sort(ord.begin(), ord.end(), [&](int a, int b2) { return ov[2 * a] < ov[2 * b2]; });
    for (int i = 0; i < ns; i++) { ob[i] = ov[2 * ord[i]]; oc[i] = ov[2 * ord[i] + 1]; }
    const TileCfg tile = default_tile(rw);
    const int64_t ptrecs = (int64_t)tile.block * tile.items;
    const int64_t ont = make_tiles(oc.data(), ns, ptrecs, ob2.data(), ot.data());
    Sizer sz;
    lay_plan(sz, ns, plain.ks.key_bytes, nullptr);
    if ((rc = c.ovf.ensure(sz.off + 4096))) return rc;
    Arena ar{(char*)c.ovf.p};
    PlanArrays pa;
    lay_plan(ar, ns, plain.ks.key_bytes, &pa);
    MPSG_CHECK(hipMemcpyAsync(pa.begin, ob.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pa.count, oc.data(), sizeof(int64_t) * ns, hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipMemcpyAsync(pa.tile_start, ot.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    if ((size_t)(ont > 0 ? ont : 1) * kBins * sizeof(uint64_t) > c.lookback.bytes)
        return fail(MPSORT_ERR_ARG, "look-back buffer too small for the bucket fallback");
    SegSortPlan op = plain;
    op.tile = tile;
    op.nseg = ns;
    op.ntiles = ont;
    op.nrec = 0;
    for (int i = 0; i < ns; i++) op.nrec += oc[i];
    op.d_begin = pa.begin;
    op.d_count = pa.count;
    op.d_tile_start = pa.tile_start;
    op.d_hist = pa.hist;
    op.d_digit_base = pa.digit_base;
    op.d_pass_flags = pa.pass_flags;
    op.d_tickets = pa.tickets;
    op.d_lookback = (uint64_t*)c.lookback.p;
    op.sparse = true;
    op.all_passes = false;
    op.d_epoch_base = nullptr;
    SortBuffers b{const_cast<void*>(src), dst, tmp};
    if ((rc = seg_count(op, src, nullptr, st))) return rc;
    if ((rc = seg_passes(op, b, st, cs_out, &passes, nullptr))) return rc;
    uint32_t e2 = 0;
    MPSG_CHECK(hipMemcpyAsync(&e2, pa.pass_flags + kMaxPasses, sizeof e2, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    if (e2) return fail(MPSORT_ERR_TIMEOUT, "bucket fallback sort: device flag 0x%x", e2);
    return 0;
}

// model-bucketed local sort of every segment of `side` from src into dst
// (tmp receives the bucket scatter).  plain is the radix plan of the same
// segments, used when too many buckets overflow.
// eo (optional): the buckets sorted in LDS also write the positions of the
// merged SecondSort's table entries.
static int learned_local_sort(DevCtx& c, int rw, const KeySpec& ks, int P, int64_t n, const LearnedSide& L,
                              const SegSortPlan& plain, const void* src, void* dst, void* tmp, long long* cs_in,
                              long long* cs_out, LearnedCounters* lc, hipStream_t st, hipEvent_t* ev,
                              const EntryOut* eo = nullptr, const int64_t* hcount = nullptr) {
    int rc = 0;
    // ev (optional): 6 events around bucket histogram, scatter, in-LDS
    // bucket sort, in-LDS radix sort of crowded buckets and the overflow
    // fallback
    auto mark = [&](int i) -> int {
        if (ev) MPSG_CHECK(hipEventRecord(ev[i], st));
        return 0;
    };
    const LKey lk = make_lkey(ks);
    const int64_t trecs = lscatter_tile(rw);
    {
        // the counters cleared by one launch (the bucket counts in pieces
        // of 8192 words, one workgroup each)
        FanOut f;
        f.n = 0;
        const int64_t nh = (int64_t)P * kLBins;
        for (int64_t o = 0; o < nh && f.n < kFanMax - 3; o += 8192)
            f.e[f.n++] = FanEntry{L.lhist + o, nullptr, (uint32_t)std::min<int64_t>(8192, nh - o), 0u};
        if (nh > (int64_t)8192 * (kFanMax - 3)) {
            f.n = 0;
            if ((rc = launch_fill_u32(L.lhist, 0u, nh, st))) return rc;
        }
        f.e[f.n++] = FanEntry{L.tickets, nullptr, 4u, 0u};
        f.e[f.n++] = FanEntry{L.err, nullptr, 4u, 0u};
        f.e[f.n++] = FanEntry{L.noverflow, nullptr, 4u, 0u};
        if ((rc = launch_fanout(f, st))) return rc;
    }
    lc->lists[lc->lists[0] ? 1 : 0] = L.noverflow;
    // tiles of the learned scatter (the look-back buffer was sized by the
    // caller: it must not move while plans hold its address)
    // (hcount: the segment sizes on the host, as uploaded to L.count)
    std::vector<int64_t> cnt(P), beg(P), lt(P + 1);
    if (hcount) {
        std::copy(hcount, hcount + P, cnt.begin());
    } else {
        MPSG_CHECK(hipMemcpyAsync(cnt.data(), L.count, sizeof(int64_t) * P, hipMemcpyDeviceToHost, st));
        MPSG_CHECK(hipStreamSynchronize(st));
    }
    {
        // segments whose average bucket would exceed half the in-LDS capacity
        // would mostly overflow: such a side takes the plain radix sort
        const int64_t big = *std::max_element(cnt.begin(), cnt.end());
        if (big > (int64_t)kLBins * kChunkCap / 2 || (g_faults & MPSORT_GPU_FAULT_PLAIN)) {
            lc->global_fallback++;
            lc->unbucketed++;
            if ((rc = mark(0))) return rc;
            int passes = 0;
            SortBuffers b{const_cast<void*>(src), dst, tmp};
            if ((rc = segmented_sort(plain, b, st, cs_in != nullptr, cs_out != nullptr, &passes, nullptr))) return rc;
            for (int e = 1; e <= 5; e++)
                if ((rc = mark(e))) return rc;
            return 0;
        }
    }
    const int64_t ntiles = make_tiles(cnt.data(), P, trecs, beg.data(), lt.data());
    // staged in the per-call pinned area when it has room (no wait for the
    // copy); otherwise copied from lt, which must outlive the copy
    if (int64_t* plt = (int64_t*)pin_take(c, sizeof(int64_t) * (P + 1))) {
        std::copy(lt.begin(), lt.end(), plt);
        MPSG_CHECK(hipMemcpyAsync(L.ltiles, plt, sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
    } else {
        MPSG_CHECK(hipMemcpyAsync(L.ltiles, lt.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipStreamSynchronize(st));
    }
    // per-tile bucket counts (u16) and per-tile offsets (u32)
    const size_t tc_bytes = ((size_t)(ntiles > 0 ? ntiles : 1) * kLBins * sizeof(uint16_t) + 255) & ~(size_t)255;
    const size_t to_bytes = ((size_t)(ntiles > 0 ? ntiles : 1) * kLBins * sizeof(uint32_t) + 255) & ~(size_t)255;
    if ((rc = c.ltile.ensure(tc_bytes + to_bytes + (size_t)n * sizeof(uint16_t)))) return rc;
    uint16_t* tile_cnt = (uint16_t*)c.ltile.p;
    uint32_t* tile_off = (uint32_t*)((char*)c.ltile.p + tc_bytes);
    uint16_t* bkt = (uint16_t*)((char*)c.ltile.p + tc_bytes + to_bytes);  // bucket of every record
    SegDev sd{L.begin, L.count, L.ltiles, P, 0};
    if ((rc = mark(0))) return rc;
    if ((rc = launch_lhist(rw, src, sd, ntiles, (int)trecs, lk, L.tables, L.tstride, L.lhist, tile_cnt, cs_in, bkt,
                           st)))
        return rc;
    if ((rc = launch_lscan(P, L.lhist, L.begin, L.bbase, st))) return rc;
    if ((rc = launch_ltile_scan(tile_cnt, sd, L.bbase, tile_off, st))) return rc;
    if ((rc = mark(1))) return rc;
    // single-value buckets are written to their final slots by the scatter
    // unless that would overwrite records the scatter still reads
    const bool eq_direct = dst != src;
    if ((rc = launch_lscatter(rw, src, tmp, eq_direct ? dst : tmp, sd, ntiles, bkt, L.tables, L.tstride, tile_off,
                              L.err, st)))
        return rc;
    if ((rc = mark(2))) return rc;
    // buckets up to 4096 records: 512-thread workgroups (two per CU); larger
    // ones (up to kChunkCap) by the 1024-thread form from a list, then the
    // crowded ones by the LDS radix sort
    if ((rc = launch_chunk_sort_small(rw, tmp, dst, lk, L.lhist, L.bbase, L.tables, L.tstride, (int64_t)P * kLBins,
                                      L.overflow, L.noverflow, kMaxOverflowList, L.hard, L.noverflow + 1, L.medium,
                                      L.noverflow + 2, cs_out, eq_direct, st, eo, fault_cap())))
        return rc;
    if ((rc = mark(3))) return rc;
    uint32_t novh[3] = {0, 0, 0}, errs = 0;
    {
        Readback rb(c, st, true);
        if ((rc = rb.add(novh, L.noverflow, sizeof novh)) || (rc = rb.add(&errs, L.err, sizeof errs)) ||
            (rc = rb.wait()))
            return rc;
    }
    if (errs) return fail(MPSORT_ERR_TIMEOUT, "learned bucket scatter: device flag 0x%x", errs);
    if ((rc = launch_listed_sort(rw, tmp, dst, lk, L.medium, novh[2], L.lhist, L.bbase, L.tables, kLBins, L.tstride,
                                 L.hard, L.noverflow + 1, cs_out, st, eo)))
        return rc;
    if ((rc = launch_lsd_sort(rw, tmp, dst, lk, L.lhist, L.bbase, L.hard, L.noverflow + 1, (int64_t)P * kLBins, cs_out,
                              st)))
        return rc;
    if ((rc = mark(4))) return rc;
    const uint32_t nov = novh[0];
    lc->overflow_buckets += (int)nov;
    if (nov == 0) return mark(5);
    const size_t recbytes = (size_t)rw * 4;
    if ((rc = c.s2.ensure((size_t)n * recbytes))) return rc;
    int passes = 0;
    if (nov > (uint32_t)kMaxFallbackSegs) {
        // too many oversized buckets: the plain radix sort of the whole side
        lc->global_fallback++;
        if (cs_out) {
            if ((rc = launch_fill_u32((uint32_t*)cs_out, 0u, 2, st))) return rc;
        }
        // src is intact unless the sort is in place; then tmp holds every
        // record (no single-value bucket went to dst)
        SortBuffers b{src == dst ? tmp : const_cast<void*>(src), dst, c.s2.p};
        if ((rc = segmented_sort(plain, b, st, false, cs_out != nullptr, &passes, nullptr))) return rc;
        return mark(5);
    }
    // radix sort of the oversized buckets only (they are segments of tmp)
    if ((rc = sort_listed_segments(c, plain, rw, L.overflow, nov, tmp, dst, c.s2.p, cs_out, st))) return rc;
    return mark(5);
}


// The histogram sort with the learned local sorts and learned splitter
// proposals (same stages and outputs as run_histsort).
// The SecondSort without an exchange (at most kMergeMaxPH senders,
// positions in a segment below 2^31): the bounds bnd of every receiver
// bucket in the senders' segments (written before, see launch_merge_bounds)
// are clamped to the receivers' windows, the buckets grouped, gathered in
// sender order from the senders' sorted segments src and sorted in LDS
// into dst; buckets above the in-LDS capacity take the 1024-thread form, the
// LDS radix sort or the radix fallback (tmp: its third buffer).  LB holds
// the receivers' tables and lists; ev (optional): 8 timing events.
static int merged_second_sort(DevCtx& c, int rw, const LKey& lk, int P, int64_t n, const LearnedSide& LB,
                              const SegSortPlan& planB, const HistArrays& h, const int64_t* seg_begin, int32_t* bnd,
                              const void* src, void* dst, void* tmp, long long* cs_out, LearnedCounters& lc,
                              hipEvent_t* ev, hipStream_t st) {
    int rc = 0;
    // no exchange: every receiver bucket is gathered from the senders'
    // sorted segments and sorted in LDS
    auto markB = [&](int i) -> int {
        if (ev) MPSG_CHECK(hipEventRecord(ev[i], st));
        return 0;
    };
    if ((rc = markB(0))) return rc;
    // the bounds, clamped to each receiver's window in each sender
    if ((rc = launch_merge_count(bnd, P, h.myC, LB.lhist, st))) return rc;
    if ((rc = launch_lscan(P, LB.lhist, h.C, LB.bbase, st, kMergeBins))) return rc;
    if ((rc = launch_fill_u32(LB.noverflow, 0u, 8, st))) return rc;
    lc.lists[1] = LB.noverflow;
    MPSG_CHECK(hipEventRecord(c.ev[kEvExchange], st));
    if ((rc = markB(1)) || (rc = markB(2))) return rc;
    // groups of neighbouring buckets, one workgroup each; buckets
    // above the group capacity alone, in the larger form
    uint32_t novh[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if ((rc = launch_merge_group(LB.lhist, P, LB.tasks, LB.noverflow + 3, st))) return rc;
    MPSG_CHECK(hipMemcpyAsync(novh, LB.noverflow, sizeof novh, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    // descriptors (groups, then the larger buckets): group, then per
    // sender sub-run start and position
    const uint32_t nsmall = novh[3], nbig = novh[4], ntasks = nsmall + nbig;
    const size_t d_bytes = ((size_t)(ntasks ? ntasks : 1) * sizeof(MergeTask) + 255) & ~(size_t)255;
    const size_t b_bytes = (size_t)(ntasks ? ntasks : 1) * 64 * sizeof(int64_t);
    if ((rc = c.mdesc.ensure(d_bytes + b_bytes + (size_t)(ntasks ? ntasks : 1) * 64 * sizeof(int32_t))))
        return rc;
    MergeTask* desc = (MergeTask*)c.mdesc.p;
    int64_t* sub_beg = (int64_t*)((char*)c.mdesc.p + d_bytes);
    int32_t* sub_pref = (int32_t*)((char*)c.mdesc.p + d_bytes + b_bytes);
    if ((rc = launch_merge_desc(LB.tasks, nsmall, bnd, seg_begin, P, LB.bbase, LB.tables, desc, sub_beg,
                                sub_pref, st)))
        return rc;
    if ((rc = launch_merge_desc(LB.tasks + (size_t)P * kMergeBins, nbig, bnd, seg_begin, P, LB.bbase,
                                LB.tables, desc + nsmall, sub_beg + (size_t)nsmall * 64,
                                sub_pref + (size_t)nsmall * 64, st)))
        return rc;
    if ((rc = markB(6))) return rc;
    for (int big = 0; big < 2; big++) {
        const uint32_t o = big ? nsmall : 0;
        if ((rc = launch_merge_bucket(rw, big != 0, src, dst, lk, seg_begin, bnd, P, LB.lhist,
                                      LB.bbase, desc + o, sub_beg + (size_t)o * 64,
                                      sub_pref + (size_t)o * 64, big ? nbig : nsmall, LB.overflow,
                                      LB.noverflow, kMaxOverflowList, LB.hard, LB.noverflow + 1, LB.medium,
                                      LB.noverflow + 2, cs_out, st,
                                      fault_cap())))
            return rc;
    }
    if ((rc = markB(7))) return rc;
    MPSG_CHECK(hipMemcpyAsync(novh, LB.noverflow, sizeof novh, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    if ((rc = markB(3))) return rc;
    // buckets above the workgroup's capacity and crowded ones were
    // copied in place in sender order: the 1024-thread in-LDS sort,
    // then the LDS radix sort
    if ((rc = launch_listed_sort(rw, dst, dst, lk, LB.medium, novh[2], LB.lhist, LB.bbase,
                                 LB.tables, kMergeBins, kMergeBounds, LB.hard, LB.noverflow + 1, nullptr, st)))
        return rc;
    if ((rc = launch_lsd_sort(rw, dst, dst, lk, LB.lhist, LB.bbase, LB.hard, LB.noverflow + 1,
                              (int64_t)P * kMergeBins, nullptr, st)))
        return rc;
    if ((rc = markB(4))) return rc;
    const uint32_t nov = novh[0];
    lc.overflow_buckets += (int)nov;
    if (nov > (uint32_t)kMaxFallbackSegs || (g_faults & MPSORT_GPU_FAULT_PLAIN)) {
        // more oversized buckets than the bucket fallback takes (the
        // list may also have stopped at its capacity): every one of
        // them was copied to its slots in sender order, so a radix
        // sort of each receiver's whole segment finishes the sort
        lc.global_fallback++;
        if (cs_out && (rc = launch_fill_u32((uint32_t*)cs_out, 0u, 2, st))) return rc;
        int passes = 0;
        SortBuffers b{dst, dst, tmp};
        if ((rc = segmented_sort(planB, b, st, false, cs_out != nullptr, &passes, nullptr))) return rc;
        // no digit pass ran (one key value): the output checksum is
        // taken by a copy
        if (passes == 0 && cs_out && (rc = launch_copy(dst, tmp, n, rw, cs_out, st)))
            return rc;
    } else if (nov && (rc = sort_listed_segments(c, planB, rw, LB.overflow, nov, dst, dst, tmp,
                                                 nullptr, st))) {
        return rc;
    }
    if ((rc = markB(5))) return rc;
    return 0;
}

// The plain path's merged SecondSort: receiver tables from an even sample
// of each receiver's records, their positions searched in the receivers'
// windows of the senders' sorted segments src (no model, no exchange).
// ev: the SecondSort's timing events (ev[8] marks the tables' start).
static int plain_merged_second(DevCtx& c, const CallShape& cs, const SegSortPlan& planB, const HistArrays& h,
                               const PlanArrays& pa, const void* src, void* dst, void* tmp, long long* cs_out,
                               hipEvent_t* ev, MergedReport* rep, hipStream_t st) {
    const int P = cs.P, rw = cs.rw;
    int rc = 0;
    // the first of the two learned-side layouts (the bucket FirstSort uses
    // the second)
    Sizer lz;
    lay_lside(lz, P, nullptr);
    lay_lside(lz, P, nullptr);
    if ((rc = c.lsmall.ensure(lz.off + 4096))) return rc;
    Arena la{(char*)c.lsmall.p};
    LearnedSide LB;
    lay_lside(la, P, &LB);
    LB.tstride = kMergeBounds;
    // bounds, then the receivers' samples and their window prefixes
    const size_t bnd_bytes = ((size_t)P * P * (kMergeBins + 1) * sizeof(int32_t) + 255) & ~(size_t)255;
    const size_t samp_bytes = (size_t)P * kMergeSamples * sizeof(uint64_t);
    if ((rc = c.mbnd.ensure(bnd_bytes + samp_bytes + (size_t)P * (kMergeMaxPH + 1) * sizeof(int64_t)))) return rc;
    int32_t* bnd = (int32_t*)c.mbnd.p;
    uint64_t* samp = (uint64_t*)((char*)c.mbnd.p + bnd_bytes);
    int64_t* spref = (int64_t*)((char*)c.mbnd.p + bnd_bytes + samp_bytes);
    const LKey lk = make_lkey(cs.ks);
    if (ev) MPSG_CHECK(hipEventRecord(ev[8], st));
    if ((rc = launch_merge_sample_tables(rw, src, lk, pa.begin, P, h.myC, LB.tables, samp, spref, st))) return rc;
    if ((rc = launch_merge_bounds(rw, src, lk, pa.begin, pa.count, P, LB.tables, nullptr, 0, nullptr, bnd, st,
                                  h.myC, samp, spref)))
        return rc;
    LearnedCounters lc;
    if ((rc = merged_second_sort(c, rw, lk, P, cs.ntot, LB, planB, h, pa.begin, bnd, src, dst, tmp, cs_out, lc, ev,
                                 st)))
        return rc;
    rep->lists_dev = LB.noverflow;
    rep->overflow = lc.overflow_buckets;
    rep->global_fallback = lc.global_fallback;
    return 0;
}

// FirstSort tables from the records: kFirstSamples keys of every segment
// (L.begin / L.count on the device) at evenly spaced positions, sorted by a
// segmented radix sort of the sample, give each segment a table of
// kLBounds boundaries in L.tables (stride kLBounds).  plan supplies the
// look-back flags and the checksum words.
static int sample_first_tables(DevCtx& c, const CallShape& cs, const SegSortPlan& plan, const LearnedSide& L,
                               const void* src, hipStream_t st, const uint32_t** flag_dev) {
    const int P = cs.P, rw = cs.rw, S = kFirstSamples;
    int rc = 0;
    const size_t sbytes = ((size_t)P * S * sizeof(uint64_t) + 255) & ~(size_t)255;
    if ((rc = c.fsamp.ensure(3 * sbytes))) return rc;
    uint64_t* s0 = (uint64_t*)c.fsamp.p;
    uint64_t* s1 = (uint64_t*)((char*)c.fsamp.p + sbytes);
    uint64_t* s2 = (uint64_t*)((char*)c.fsamp.p + 2 * sbytes);
    KeySpec sk;
    fill_keyspec(&sk, 2, 0, 8, 8, 0);
    std::vector<int64_t> sc(P, S), sb(P), st_(P + 1);
    const TileCfg stile = default_tile(2);
    const int64_t snt = make_tiles(sc.data(), P, (int64_t)stile.block * stile.items, sb.data(), st_.data());
    if ((size_t)snt * kBins * sizeof(uint64_t) > c.lookback.bytes)
        return fail(MPSORT_ERR_ARG, "look-back buffer too small for the FirstSort sample");
    // the sample's segments depend on P only: their arrays stay on the
    // device from one call to the next
    Sizer pz;
    lay_plan(pz, P, 8, nullptr);
    const void* was = c.fsplan.p;
    if ((rc = c.fsplan.ensure(pz.off + 4096))) return rc;
    Arena pr{(char*)c.fsplan.p};
    PlanArrays spa;
    lay_plan(pr, P, 8, &spa);
    const int64_t key = ((int64_t)P << 32) | ((int64_t)stile.block << 8) | stile.items;
    if (c.fsplan.p != was || c.fsplan_key != key) {
        c.fsplan_key = 0;
        MPSG_CHECK(hipMemcpyAsync(spa.begin, sb.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(spa.count, sc.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(spa.tile_start, st_.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipStreamSynchronize(st));
        c.fsplan_key = key;
    }
    if ((rc = launch_fill_u32(spa.pass_flags, 0u, kMaxPasses + 1, st))) return rc;
    SegSortPlan sp = plan;
    sp.rw = 2;
    sp.tile = stile;
    sp.ks = sk;
    sp.nrec = (int64_t)P * S;
    sp.nseg = P;
    sp.ntiles = snt;
    sp.d_begin = spa.begin;
    sp.d_count = spa.count;
    sp.d_tile_start = spa.tile_start;
    sp.d_hist = spa.hist;
    sp.d_digit_base = spa.digit_base;
    sp.d_pass_flags = spa.pass_flags;
    sp.d_tickets = spa.tickets;
    sp.sparse = false;
    sp.all_passes = false;
    sp.d_epoch_base = nullptr;
    const LKey lk = make_lkey(cs.ks);
    if ((rc = launch_seg_sample(rw, src, lk, L.begin, L.count, P, S, s0, st))) return rc;
    int passes = 0;
    SortBuffers sbuf{s0, s1, s2};
    if ((rc = segmented_sort(sp, sbuf, st, false, false, &passes, nullptr))) return rc;
    // no pass ran: every sample key is equal and the sample stayed in s0
    if ((rc = launch_sample_tables(passes ? s1 : s0, P, S, L.tables, st))) return rc;
    // the sample sort's error flag is read at the call's end (whatever the
    // tables, the bucket sort's result is the sorted order)
    *flag_dev = spa.pass_flags + kMaxPasses;
    return 0;
}

// The plain path's bucket FirstSort: tables from a sample of every segment
// (sample_first_tables), then the learned path's local sort (bucket
// histogram, scatter, in-LDS bucket sorts, fallbacks) from src into dst.
// hbegin / hcount: the segments on the host.
static int plain_bucket_first(DevCtx& c, const CallShape& cs, const SegSortPlan& planA, const int64_t* hbegin,
                              const int64_t* hcount, const void* src, void* dst, void* tmp, long long* cs_in,
                              long long* cs_out, hipEvent_t* ev, MergedReport* rep, hipStream_t st) {
    const int P = cs.P;
    int rc = 0;
    Sizer lz;
    lay_lside(lz, P, nullptr);
    lay_lside(lz, P, nullptr);
    if ((rc = c.lsmall.ensure(lz.off + 4096))) return rc;
    Arena la{(char*)c.lsmall.p};
    LearnedSide LM, LA;
    lay_lside(la, P, &LM);  // the merged SecondSort's
    lay_lside(la, P, &LA);
    LA.tstride = kLBounds;
    {
        // hbegin / hcount lie in the pinned table buffer
        Upload up(c, st, c.pinned, c.pinned_bytes);
        if ((rc = up.add(LA.begin, hbegin, sizeof(int64_t) * P)) || (rc = up.add(LA.count, hcount, sizeof(int64_t) * P)) ||
            (rc = up.flush()))
            return rc;
    }
    if (ev) MPSG_CHECK(hipEventRecord(ev[8], st));
    if ((rc = sample_first_tables(c, cs, planA, LA, src, st, &rep->flag_dev))) return rc;
    LearnedCounters lc;
    if ((rc = learned_local_sort(c, cs.rw, cs.ks, P, cs.ntot, LA, planA, src, dst, tmp, cs_in, cs_out, &lc, st, ev,
                                 nullptr, hcount)))
        return rc;
    rep->lists_dev = LA.noverflow;
    rep->overflow = lc.overflow_buckets;
    rep->global_fallback = lc.global_fallback;
    return 0;
}

static int run_histsort_learned(DevCtx& c, const CallShape& cs, const void* d_in, void* d_out,
                                const int64_t* nin, const int64_t* nout, const LearnedModel& M,
                                mpsort_gpu_stats* stats) {
    hipStream_t st = c.st;
    const int P = cs.P;
    const int rw = cs.rw;
    const int64_t n = cs.ntot;
    const size_t recbytes = (size_t)rw * 4;
    const int npass = cs.ks.key_bytes;
    int rc = 0;

    Sizer sz;
    lay_plan(sz, P, npass, nullptr);
    lay_plan(sz, P, npass, nullptr);
    lay_hist(sz, P, nullptr);
    if ((rc = c.small.ensure(sz.off + 4096))) return rc;
    Arena ar{(char*)c.small.p};
    PlanArrays pa, pb;
    HistArrays h;
    lay_plan(ar, P, npass, &pa);
    lay_plan(ar, P, npass, &pb);
    lay_hist(ar, P, &h);
    Sizer lz;
    lay_lside(lz, P, nullptr);
    lay_lside(lz, P, nullptr);
    lz.take<int>((size_t)P + 1);
    if ((rc = c.lsmall.ensure(lz.off + 4096))) return rc;
    Arena la{(char*)c.lsmall.p};
    LearnedSide LA, LB;
    lay_lside(la, P, &LA);
    lay_lside(la, P, &LB);
    int* exact = la.take<int>((size_t)P + 1);

    if ((rc = c.s0.ensure(n * recbytes))) return rc;
    if ((rc = c.s1.ensure(n * recbytes))) return rc;
    if (cs.general && (rc = c.rbuf.ensure(n * recbytes))) return rc;

    const size_t tab = (size_t)(9 * P + 4);
    if ((rc = ensure_pinned(c, tab * sizeof(int64_t) + 4096))) return rc;
    int64_t* ht = (int64_t*)c.pinned;
    int64_t *hbA = ht, *hcA = ht + P, *htA = ht + 2 * P;
    int64_t *hbB = ht + 3 * P + 1, *hcB = ht + 4 * P + 1, *htB = ht + 5 * P + 1;
    int64_t* hC = ht + 6 * P + 2;
    float* hq = (float*)(ht + 7 * P + 3);  // 4 * P floats
    for (int t = 0; t < P; t++) { hcA[t] = nin[t]; hcB[t] = nout[t]; }
    const TileCfg tile = default_tile(rw);
    const int64_t trecs = (int64_t)tile.block * tile.items;
    const int64_t ntA = make_tiles(hcA, P, trecs, hbA, htA);
    const int64_t ntB = make_tiles(hcB, P, trecs, hbB, htB);
    hC[0] = 0;
    for (int t = 0; t < P; t++) hC[t + 1] = hC[t] + nout[t];
    for (int t = 0; t < P; t++) {
        hq[t] = 0.f;
        hq[P + t] = 1.f;
        hq[2 * P + t] = (float)((double)hC[t] / (double)n);
        hq[3 * P + t] = (float)((double)hC[t + 1] / (double)n);
    }
    const int64_t ntiles = ntA > ntB ? ntA : ntB;
    {
        // one size for every radix pass of this call (whole sides and the
        // bucket fallback)
        const size_t need = (size_t)(ntiles + kMaxFallbackSegs + 1) * kBins * sizeof(uint64_t);
        if ((rc = c.lookback.ensure(need))) return rc;
    }
    if ((rc = ensure_rpin(c, (size_t)64 * (P + 16) * sizeof(int64_t)))) return rc;
    {
        Upload up(c, st, c.pinned, c.pinned_bytes);
        if ((rc = up.add(pa.begin, hbA, sizeof(int64_t) * P)) || (rc = up.add(pa.count, hcA, sizeof(int64_t) * P)) ||
            (rc = up.add(pa.tile_start, htA, sizeof(int64_t) * (P + 1))) ||
            (rc = up.add(pb.begin, hbB, sizeof(int64_t) * P)) || (rc = up.add(pb.count, hcB, sizeof(int64_t) * P)) ||
            (rc = up.add(pb.tile_start, htB, sizeof(int64_t) * (P + 1))) ||
            (rc = up.add(h.C, hC, sizeof(int64_t) * (P + 1))) || (rc = up.add(LA.begin, hbA, sizeof(int64_t) * P)) ||
            (rc = up.add(LA.count, hcA, sizeof(int64_t) * P)) || (rc = up.add(LB.begin, hbB, sizeof(int64_t) * P)) ||
            (rc = up.add(LB.count, hcB, sizeof(int64_t) * P)) || (rc = up.add(LA.q0, hq, sizeof(float) * P)) ||
            (rc = up.add(LA.q1, hq + P, sizeof(float) * P)) || (rc = up.add(LB.q0, hq + 2 * P, sizeof(float) * P)) ||
            (rc = up.add(LB.q1, hq + 3 * P, sizeof(float) * P)) || (rc = up.fill(h.err, 0u, 4)) ||
            (rc = up.fill((uint32_t*)h.checksum, 0u, 4)) || (rc = up.fill(pa.pass_flags, 0u, kMaxPasses + 1)) ||
            (rc = up.fill(pb.pass_flags, 0u, kMaxPasses + 1)) || (rc = up.flush()))
            return rc;
    }

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
    SegSortPlan planA = mkplan(pa, ntA);
    SegSortPlan planB = mkplan(pb, ntB);
    const ModelDev md = M.dev();

    MPSG_CHECK(hipEventRecord(c.ev[kEvStart], st));
    void* recs_in = const_cast<void*>(d_in);
    void* final_dst = d_out;
    if (cs.general) {
        if ((rc = launch_pack(d_in, n, cs.elsize, cs.key_off_user, cs.key_bytes_user, c.rbuf.p, rw, cs.iwords,
                              cs.checksum ? h.checksum : nullptr, st)))
            return rc;
        recs_in = c.rbuf.p;
        final_dst = c.rbuf.p;
    }
    const bool cs_direct = cs.checksum && !cs.general;
    LearnedCounters lc;
    hipEvent_t* evA = cs.timing ? c.pass_ev[0] : nullptr;
    hipEvent_t* evB = cs.timing && P > 1 ? c.pass_ev[1] : nullptr;
    bool merged_side = false;
    int overflow_first = 0;

    // FirstSort: one table per segment over the quantile range its keys
    // cover (the whole distribution unless a sample of the segment's keys
    // says they come from one part of it, as in nearly sorted input)
    LA.tstride = kLBounds;
    LB.tstride = kLBounds;
    {
        uint64_t xm = 0;
        for (int b = 0; b < cs.ks.key_bytes && b < 8; b++) xm |= (uint64_t)cs.ks.xmask[b] << (8 * b);
        if ((rc = launch_seg_range(md, recs_in, (int64_t)recbytes, cs.ks.key_off, cs.ks.key_bytes, xm, LA.begin,
                                   LA.count, P, LA.q0, LA.q1, LA.nedge, st)))
            return rc;
    }
    if ((rc = launch_model_tables(md, M.kmax, LA.q0, LA.q1, P, kLBounds, LA.tables, M.heavy, M.heavy_q, M.nheavy, st,
                                  LA.nedge)))
        return rc;
    if (P == 1) {
        if ((rc = learned_local_sort(c, rw, cs.ks, P, n, LA, planA, recs_in, final_dst, c.s1.p,
                                     cs_direct ? h.checksum : nullptr, cs_direct ? h.checksum + 1 : nullptr, &lc,
                                     st, evA, nullptr, hcA)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        for (int e = kEvFindP; e <= kEvSecond; e++) MPSG_CHECK(hipEventRecord(c.ev[e], st));
    } else {
        // SecondSort: receiver r holds the quantile range [C[r], C[r+1]) / n;
        // the merged form uses kMergeBins buckets per receiver (positions
        // inside a sender's segment are kept as int32)
        const bool merged = merge_eligible(cs);
        merged_side = merged;
        const int nbB = merged ? kMergeBounds : kLBounds;
        LB.tstride = nbB;
        const LKey lkB = make_lkey(cs.ks);
        int32_t* bnd = nullptr;
        EntryOut eo{};
        if (merged) {
            // the receivers' tables need only the model and C: made now, so
            // that the FirstSort's bucket sorts write the position of every
            // table entry in their sorted buckets (no search afterwards)
            const size_t bnd_bytes = ((size_t)P * P * (kMergeBins + 1) * sizeof(int32_t) + 255) & ~(size_t)255;
            const size_t rng_bytes = ((size_t)P * (kLBins + 1) * sizeof(uint32_t) + 255) & ~(size_t)255;
            if ((rc = c.mbnd.ensure(bnd_bytes + rng_bytes + (size_t)P * kLBins))) return rc;
            bnd = (int32_t*)c.mbnd.p;
            if ((rc = launch_model_tables(md, M.kmax, LB.q0, LB.q1, P, nbB, LB.tables, M.heavy, M.heavy_q, M.nheavy, st, nullptr,
                                          true)))
                return rc;
            // the edge entries are placed once the windows are known
            if ((rc = launch_tables_reserve_edges(LB.tables, P, st))) return rc;
            eo.keys = LB.tables;
            eo.range = (uint32_t*)((char*)c.mbnd.p + bnd_bytes);
            eo.seg_begin = pa.begin;
            eo.bnd = bnd;
            eo.done = (uint8_t*)((char*)c.mbnd.p + bnd_bytes + rng_bytes);
            eo.P = P;
            if ((rc = launch_entry_ranges(LA.tables, LA.tstride, P, LB.tables, const_cast<uint32_t*>(eo.range), st)))
                return rc;
            if ((rc = launch_fill_u32((uint32_t*)eo.done, 0u, (int64_t)P * kLBins / 4, st))) return rc;
        }
        if ((rc = learned_local_sort(c, rw, cs.ks, P, n, LA, planA, recs_in, c.s0.p, c.s1.p,
                                     cs_direct ? h.checksum : nullptr, nullptr, &lc, st, evA, merged ? &eo : nullptr,
                                     hcA)))
            return rc;
        overflow_first = lc.overflow_buckets;
        if (merged) {
            if (lc.unbucketed == 0) {
                // entries of the buckets sorted outside LDS, and the fixed ends
                if ((rc = launch_entries_fixup(rw, c.s0.p, lkB, eo, LA.lhist, LA.bbase, pa.count, st))) return rc;
            } else {
                // the side was radix-sorted without buckets: search every
                // entry in every sender's whole segment
                if ((rc = launch_merge_bounds(rw, c.s0.p, lkB, pa.begin, pa.count, P, LB.tables, nullptr, 0, nullptr,
                                              bnd, st)))
                    return rc;
            }
        }
        MPSG_CHECK(hipEventRecord(c.ev[kEvFirst], st));
        // splitters: values from the model, each checked with exact counts
        if ((rc = launch_bisect(c.s0.p, rw, cs.ks, pa.begin, pa.count, P, h.C, h.win, h.myclt, h.mycle, h.split,
                                h.pminmax, h.rounds, h.err, st, &md, exact,
                                (g_faults & MPSORT_GPU_FAULT_SPLITTERS) != 0)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvFindP], st));
        if ((rc = launch_layout(P, h.C, pa.count, pa.begin, h.myclt, h.mycle, h.myC, h.send,
                                exch_runs(h, P, cs.exch_policy), h.err, st)))
            return rc;
        MPSG_CHECK(hipEventRecord(c.ev[kEvLayout], st));
        if (!merged &&
            (rc = launch_model_tables(md, M.kmax, LB.q0, LB.q1, P, nbB, LB.tables, M.heavy, M.heavy_q, M.nheavy, st)))
            return rc;
        if (merged) {
            if ((rc = launch_merge_edges(rw, c.s0.p, lkB, pa.begin, P, h.myC, LB.tables, bnd, st))) return rc;
            if ((rc = merged_second_sort(c, rw, lkB, P, n, LB, planB, h, pa.begin, bnd, c.s0.p, final_dst, c.s1.p,
                                         cs_direct ? h.checksum + 1 : nullptr, lc, evB, st)))
                return rc;
        } else {
            if ((rc = launch_exchange(c.s0.p, c.s1.p, n, rw, run_table(h, P), h.err, st))) return rc;
            MPSG_CHECK(hipEventRecord(c.ev[kEvExchange], st));
            if ((rc = learned_local_sort(c, rw, cs.ks, P, n, LB, planB, c.s1.p, final_dst, c.s0.p, nullptr,
                                         cs_direct ? h.checksum + 1 : nullptr, &lc, st, evB, nullptr, hcB)))
                return rc;
        }
        MPSG_CHECK(hipEventRecord(c.ev[kEvSecond], st));
    }
    if (cs.general) {
        if ((rc = launch_gather(d_in, n, cs.elsize, c.rbuf.p, rw, cs.iwords, d_out,
                                cs.checksum ? h.checksum + 1 : nullptr, h.err, st)))
            return rc;
    }
    if ((g_faults & MPSORT_GPU_FAULT_CORRUPT) && (rc = launch_fault_corrupt(nullptr, h.checksum, st))) return rc;
    MPSG_CHECK(hipEventRecord(c.ev[kEvEnd], st));

    uint32_t err0 = 0;
    long long sums[2] = {0, 0};
    std::vector<int> rounds(P > 1 ? P - 1 : 1, 0), ex(P > 1 ? P - 1 : 1, 0);
    // list counters of each side: [1] LDS radix sort, [2] 1024-thread form
    uint32_t lists[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int sel[2] = {0, 0};
    {
        Readback rb(c, st, true);
        for (int side = 0; side < 2; side++)
            if (lc.lists[side] && (rc = rb.add(lists[side], lc.lists[side], sizeof lists[side]))) return rc;
        if ((rc = rb.add(&err0, h.err, sizeof err0)) || (rc = rb.add(sums, h.checksum, sizeof sums))) return rc;
        if (P > 1 && ((rc = rb.add(rounds.data(), h.rounds, sizeof(int) * (P - 1))) ||
                      (rc = rb.add(ex.data(), exact, sizeof(int) * (P - 1))) || (rc = rb.add(sel, h.sel, sizeof sel))))
            return rc;
        if ((rc = rb.wait())) return rc;
    }
    int rmax = 0, nexact = 0;
    for (int i = 0; i + 1 < P; i++) {
        rmax = std::max(rmax, rounds[i]);
        nexact += ex[i];
    }
    stats->nvranks = P;
    stats->record_bytes = (int32_t)recbytes;
    stats->bisect_rounds = rmax;
    stats->repair_rounds = rmax;
    stats->general_path = cs.general ? 1 : 0;
    stats->nrecords = n;
    stats->checksum_in = sums[0];
    stats->checksum_out = sums[1];
    stats->learned = 1;
    stats->splitters_exact = nexact;
    stats->splitters_repaired = P > 1 ? P - 1 - nexact : 0;
    stats->buckets = P * kLBins + (P > 1 ? P * (LB.tstride == kMergeBounds ? kMergeBins : kLBins) : 0);
    stats->buckets_overflow = lc.overflow_buckets;
    stats->buckets_overflow_first = P == 1 ? lc.overflow_buckets : overflow_first;
    stats->t_first_sort_ms = ev_ms(c.ev[kEvStart], c.ev[kEvFirst]);
    stats->t_findp_ms = ev_ms(c.ev[kEvFirst], c.ev[kEvFindP]);
    stats->t_laysolve_ms = ev_ms(c.ev[kEvFindP], c.ev[kEvLayout]);
    stats->t_exchange_ms = ev_ms(c.ev[kEvLayout], c.ev[kEvExchange]);
    stats->t_second_sort_ms = ev_ms(c.ev[kEvExchange], c.ev[kEvSecond]);
    stats->t_total_ms = ev_ms(c.ev[kEvStart], c.ev[kEvEnd]);
    stats->plain_sides = lc.global_fallback;
    stats->buckets_lds_radix = (int32_t)(lists[0][1] + lists[1][1]);
    stats->buckets_listed = (int32_t)(lists[0][2] + lists[1][2]);
    // the merged SecondSort gathers from the senders' runs: no exchange
    stats->exchange_sparse = P > 1 && !merged_side ? sel[0] : 0;
    stats->exchange_runs = P > 1 && !merged_side ? (sel[0] ? sel[1] : P * P) : 0;
    // pass_ms: bucket histogram, scatter, in-LDS bucket sort, in-LDS radix
    // sort, overflow fallback for the FirstSort and then the SecondSort
    int np = 0;
    for (int side = 0; side < 2; side++) {
        hipEvent_t* e = side == 0 ? evA : evB;
        if (!e) continue;
        for (int k = 0; k < 5; k++) stats->pass_ms[np++] = ev_ms(e[k], e[k + 1]);
    }
    // merged SecondSort: its gather-and-sort launch, inside the in-LDS
    // bucket sort entry
    if (evB && merged_side) stats->pass_ms[np++] = ev_ms(evB[6], evB[7]);
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
    if (err0 & 1u) return fail(MPSORT_ERR_TIMEOUT, "a look-back wait exceeded its spin limit");
    if (err0) return fail(MPSORT_ERR_LAYOUT, "layout check failed (device flags 0x%x)", err0);
    if (cs.checksum && sums[0] != sums[1])
        return fail(MPSORT_ERR_CHECKSUM, "MPSort: Data changed after sorting; checksum mismatch (%lld != %lld)",
                    sums[0], sums[1]);
    return 0;
}

}  // namespace mpsg

using namespace mpsg;

// ---- batched segmented sort ------------------------------------------------
// Thousands of independent stable sorts of one record type, segments back to
// back in one buffer.  Segments of at most `threshold` records are sorted by
// one workgroup each in LDS (the bucket-sort kernels of the learned path, with
// the segment as the bucket); the larger ones by the segmented onesweep over
// just those segments.  Every launch of a group depends only on device data,
// so the whole batch can be captured once in a hipGraph and replayed.
// tile: the onesweep tile of the large segments; threshold: the largest
// segment sorted in LDS; form: the workgroup shape of those, which sets how
// many share a CU (0: 1024 x 9, one per CU, 16 waves; 1: 512 x 9, two per
// CU, 16 waves; 2: 512 x 6, three per CU, 24 waves), the segments above its
// capacity going to the 1024-thread form
struct SegTune {
    TileCfg tile;
    int threshold;
    int form;
};
static int form_waves_per_cu(int form) { return form == 2 ? 24 : 16; }

struct SegGroup {
    const void* in = nullptr;
    void* out = nullptr;
    size_t elsize = 0;
    int rw = 0;
    KeySpec ks;
    LKey lk;
    int nseg = 0;
    int64_t nrec = 0;
    std::vector<int64_t> count, begin;
    SegTune tune{{1024, 8}, kChunkCap, 0};
    bool tuned = false;
    // device side
    Buf small;     // small-segment counts, bases, hard list, counters; epoch
    Buf plan_buf;  // arrays of the large-segment plan
    Buf tmp;       // record buffer for the digit passes
    Buf lookback;
    int nsmall = 0, nlarge = 0;
    int nsmall_form = 0;  // of the small segments, those sorted by the form chosen (the others: 1024 threads)
    uint32_t* d_cnt = nullptr;
    int64_t* d_base = nullptr;
    uint32_t* d_hard = nullptr;
    uint32_t* d_counters = nullptr;  // [0] oversized, [1] hard, [2] epoch, [4] hard of the 1024-thread list
    int64_t* d_ovf = nullptr;
    SegSortPlan plan;
    uint64_t epoch_used = 0;  // epoch steps consumed since the last reset
    // one segment at a time (MPSORT_GPU_SEGPLAN_EACH): the large segments'
    // counts and tile counts, and a {0, tiles} table per large segment
    bool each = false;
    std::vector<int64_t> lcount, ltiles;
    int64_t* d_tsv = nullptr;
};

static std::map<std::string, SegTune>& seg_tune_cache() {
    static std::map<std::string, SegTune> m;
    return m;
}

// size class: log2 of the median segment length
static std::string seg_tune_key(const SegGroup& G) {
    std::vector<int64_t> c(G.count);
    std::sort(c.begin(), c.end());
    const int64_t med = c.empty() ? 0 : c[c.size() / 2];
    int cls = 0;
    while ((int64_t(1) << (cls + 1)) <= med) cls++;
    char k[96];
    snprintf(k, sizeof k, "rw=%d keybytes=%d class=%d", G.rw, G.ks.key_bytes, cls);
    return k;
}

static std::string seg_tune_file() {
    const char* e = getenv("MPSORT_GPU_TUNE_CACHE");
    if (e) return e;
    const char* h = getenv("HOME");
    return h ? std::string(h) + "/.cache/mpsort_gpu_segsort.tune" : std::string();
}

static void seg_tune_load() {
    static bool loaded = false;
    if (loaded) return;
    loaded = true;
    const std::string f = seg_tune_file();
    if (f.empty()) return;
    FILE* fp = fopen(f.c_str(), "r");
    if (!fp) return;
    char line[256];
    while (fgets(line, sizeof line, fp)) {
        int rw, kb, cls, blk, items, thr, form = 0;
        const int got = sscanf(line, "rw=%d keybytes=%d class=%d block=%d items=%d threshold=%d form=%d", &rw, &kb,
                               &cls, &blk, &items, &thr, &form);
        if (got >= 6) {
            char k[96];
            snprintf(k, sizeof k, "rw=%d keybytes=%d class=%d", rw, kb, cls);
            SegTune t{{blk, items}, thr, form};
            if (tile_supported(rw, t.tile) && thr >= 0 && thr <= kChunkCap && form >= 0 && form <= 2)
                seg_tune_cache()[k] = t;
        }
    }
    fclose(fp);
}

static void seg_tune_store() {
    const std::string f = seg_tune_file();
    if (f.empty()) return;
    FILE* fp = fopen(f.c_str(), "w");
    if (!fp) return;
    for (auto& kv : seg_tune_cache())
        fprintf(fp, "%s block=%d items=%d threshold=%d form=%d\n", kv.first.c_str(), kv.second.tile.block,
                kv.second.tile.items, kv.second.threshold, kv.second.form);
    fclose(fp);
}

// lay out the group's device arrays for its current tuning
static int seg_group_prepare(SegGroup& G, hipStream_t st) {
    int rc = 0;
    const int npass = G.ks.key_bytes;
    const bool lds_ok = (G.rw == 2 || G.rw == 4) && G.ks.key_bytes <= 8;
    const int thr = lds_ok ? G.tune.threshold : 0;
    const int fcap = segments_lds_cap(G.tune.form);
    std::vector<uint32_t> scnt, mcnt;
    std::vector<int64_t> sbase, mbase, lcnt, lbeg;
    for (int i = 0; i < G.nseg; i++) {
        if (G.count[i] == 0) continue;
        if (G.count[i] <= thr && G.count[i] <= fcap) {
            scnt.push_back((uint32_t)G.count[i]);
            sbase.push_back(G.begin[i]);
        } else if (G.count[i] <= thr) {
            mcnt.push_back((uint32_t)G.count[i]);
            mbase.push_back(G.begin[i]);
        } else {
            lcnt.push_back(G.count[i]);
            lbeg.push_back(G.begin[i]);
        }
    }
    // the form's segments, then those above its capacity
    G.nsmall_form = (int)scnt.size();
    scnt.insert(scnt.end(), mcnt.begin(), mcnt.end());
    sbase.insert(sbase.end(), mbase.begin(), mbase.end());
    G.nsmall = (int)scnt.size();
    G.nlarge = (int)lcnt.size();
    Sizer sz;
    sz.take<uint32_t>((size_t)G.nsmall + 1);
    sz.take<int64_t>((size_t)G.nsmall + 1);
    sz.take<uint32_t>((size_t)G.nsmall + 1);
    sz.take<uint32_t>(8);
    sz.take<int64_t>(2);
    if ((rc = G.small.ensure(sz.off + 4096))) return rc;
    Arena a{(char*)G.small.p};
    G.d_cnt = a.take<uint32_t>((size_t)G.nsmall + 1);
    G.d_base = a.take<int64_t>((size_t)G.nsmall + 1);
    G.d_hard = a.take<uint32_t>((size_t)G.nsmall + 1);
    G.d_counters = a.take<uint32_t>(8);
    G.d_ovf = a.take<int64_t>(2);
    if (G.nsmall) {
        MPSG_CHECK(hipMemcpyAsync(G.d_cnt, scnt.data(), sizeof(uint32_t) * G.nsmall, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(G.d_base, sbase.data(), sizeof(int64_t) * G.nsmall, hipMemcpyHostToDevice, st));
    }
    if ((rc = launch_fill_u32(G.d_counters, 0u, 8, st))) return rc;
    G.epoch_used = 0;

    SegSortPlan& p = G.plan;
    p = SegSortPlan();
    p.rw = G.rw;
    p.tile = G.tune.tile;
    p.ks = G.ks;
    p.nseg = G.nlarge;
    p.nrec = 0;
    p.ntiles = 0;
    if (G.nlarge) {
        const int64_t T = (int64_t)p.tile.block * p.tile.items;
        std::vector<int64_t> ts(G.nlarge + 1);
        int64_t t = 0;
        for (int i = 0; i < G.nlarge; i++) {
            ts[i] = t;
            t += (lcnt[i] + T - 1) / T;
            p.nrec += lcnt[i];
        }
        ts[G.nlarge] = t;
        p.ntiles = t;
        Sizer pz;
        lay_plan(pz, G.nlarge, npass, nullptr);
        pz.take<int64_t>(2 * (size_t)G.nlarge);
        if ((rc = G.plan_buf.ensure(pz.off + 4096))) return rc;
        Arena pa_a{(char*)G.plan_buf.p};
        PlanArrays pa;
        lay_plan(pa_a, G.nlarge, npass, &pa);
        G.d_tsv = pa_a.take<int64_t>(2 * (size_t)G.nlarge);
        G.lcount = lcnt;
        G.ltiles.assign(G.nlarge, 0);
        std::vector<int64_t> tsv(2 * (size_t)G.nlarge, 0);
        for (int i = 0; i < G.nlarge; i++) G.ltiles[i] = tsv[2 * i + 1] = ts[i + 1] - ts[i];
        MPSG_CHECK(hipMemcpyAsync(G.d_tsv, tsv.data(), sizeof(int64_t) * 2 * G.nlarge, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(pa.begin, lbeg.data(), sizeof(int64_t) * G.nlarge, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(pa.count, lcnt.data(), sizeof(int64_t) * G.nlarge, hipMemcpyHostToDevice, st));
        MPSG_CHECK(hipMemcpyAsync(pa.tile_start, ts.data(), sizeof(int64_t) * (G.nlarge + 1), hipMemcpyHostToDevice,
                                  st));
        if ((rc = G.tmp.ensure((size_t)G.nrec * G.elsize))) return rc;
        G.lookback.zero_on_grow = true;
        if ((rc = G.lookback.ensure((size_t)(t + 1) * kBins * sizeof(uint64_t)))) return rc;
        MPSG_CHECK(hipMemsetAsync(G.lookback.p, 0, G.lookback.bytes, st));
        p.d_begin = pa.begin;
        p.d_count = pa.count;
        p.d_tile_start = pa.tile_start;
        p.d_hist = pa.hist;
        p.d_digit_base = pa.digit_base;
        p.d_pass_flags = pa.pass_flags;
        p.d_tickets = pa.tickets;
        p.d_lookback = (uint64_t*)G.lookback.p;
        p.d_checksum = nullptr;
        p.sparse = true;
        p.all_passes = true;
        p.d_epoch_base = G.d_counters + 2;
    }
    MPSG_CHECK(hipStreamSynchronize(st));
    return 0;
}

// the device work of one group; records nothing but kernels, so it can be
// captured (look-back epochs come from the counter bumped here)
static int seg_group_enqueue_each(SegGroup& G, hipStream_t st);
static int seg_group_enqueue(SegGroup& G, hipStream_t st) {
    int rc = 0;
    if ((rc = launch_fill_u32(G.d_counters, 0u, 2, st))) return rc;
    if ((rc = launch_fill_u32(G.d_counters + 4, 0u, 1, st))) return rc;
    if (G.each) return seg_group_enqueue_each(G, st);
    if (G.nsmall) {
        if ((rc = launch_segments_lds(G.rw, G.tune.form, G.in, G.out, G.lk, G.d_cnt, G.d_base, G.nsmall_form,
                                      G.nsmall - G.nsmall_form, G.d_ovf, G.d_counters, G.d_hard, G.d_counters + 1,
                                      G.d_counters + 4, st)))
            return rc;
    }
    if (G.nlarge) {
        if ((rc = launch_add_u32(G.d_counters + 2, (uint32_t)kMaxPasses, st))) return rc;
        if ((rc = seg_count(G.plan, G.in, nullptr, st)))
            return rc < 0 && !g_err[0] ? fail(MPSORT_ERR_UNSUPPORTED, "segmented sort: tile %d x %d not supported for "
                                              "%d-byte records", G.plan.tile.block, G.plan.tile.items, 4 * G.rw) : rc;
        SortBuffers b{const_cast<void*>(G.in), G.out, G.tmp.p};
        int passes = 0;
        if ((rc = seg_passes(G.plan, b, st, nullptr, &passes, nullptr)))
            return rc < 0 && !g_err[0] ? fail(MPSORT_ERR_UNSUPPORTED, "segmented sort: tile %d x %d not supported for "
                                              "%d-byte records", G.plan.tile.block, G.plan.tile.items, 4 * G.rw) : rc;
    }
    return 0;
}

// The same work one segment at a time: each small segment by its own
// in-LDS launch (its crowded-segment count cleared before), each large one
// by a one-segment view of the plan (its counts, histograms and digit bases;
// a {0, tiles} tile table; look-back epochs advanced per segment), all
// enqueued back to back.
static int seg_group_enqueue_each(SegGroup& G, hipStream_t st) {
    int rc = 0;
    for (int i = 0; i < G.nsmall; i++) {
        if ((rc = launch_fill_u32(G.d_counters + 1, 0u, 1, st))) return rc;
        if ((rc = launch_fill_u32(G.d_counters + 4, 0u, 1, st))) return rc;
        const bool in_form = i < G.nsmall_form;
        if ((rc = launch_segments_lds(G.rw, G.tune.form, G.in, G.out, G.lk, G.d_cnt + i, G.d_base + i, in_form ? 1 : 0,
                                      in_form ? 0 : 1, G.d_ovf, G.d_counters, G.d_hard, G.d_counters + 1,
                                      G.d_counters + 4, st)))
            return rc;
    }
    const int npass = G.ks.key_bytes;
    for (int i = 0; i < G.nlarge; i++) {
        if ((rc = launch_add_u32(G.d_counters + 2, (uint32_t)kMaxPasses, st))) return rc;
        SegSortPlan v = G.plan;
        v.nseg = 1;
        v.nrec = G.lcount[i];
        v.ntiles = G.ltiles[i];
        v.d_begin = G.plan.d_begin + i;
        v.d_count = G.plan.d_count + i;
        v.d_tile_start = G.d_tsv + 2 * i;
        v.d_hist = G.plan.d_hist + (int64_t)i * npass * kBins;
        v.d_digit_base = G.plan.d_digit_base + (int64_t)i * npass * kBins;
        if ((rc = seg_count(v, G.in, nullptr, st))) return rc < 0 && !g_err[0] ? fail(MPSORT_ERR_UNSUPPORTED, "segmented sort: tile not supported") : rc;
        SortBuffers b{const_cast<void*>(G.in), G.out, G.tmp.p};
        int passes = 0;
        if ((rc = seg_passes(v, b, st, nullptr, &passes, nullptr)))
            return rc < 0 && !g_err[0] ? fail(MPSORT_ERR_UNSUPPORTED, "segmented sort: tile not supported") : rc;
    }
    return 0;
}

// look-back epochs must stay below 2^23: reset the flags and the counter
// (outside any capture) before a run that could pass the limit
static int seg_group_epoch_guard(SegGroup& G, hipStream_t st) {
    if (!G.nlarge) return 0;
    const uint64_t step = (uint64_t)kMaxPasses * (G.each ? (uint64_t)G.nlarge + 1 : 1);
    if (G.epoch_used + 2 * step < (1u << 23)) {
        G.epoch_used += step;
        return 0;
    }
    int rc = launch_fill_u32((uint32_t*)G.lookback.p, 0u, (int64_t)(G.lookback.bytes / 4), st);
    if (rc) return rc;
    if ((rc = launch_fill_u32(G.d_counters + 2, 0u, 1, st))) return rc;
    G.epoch_used = step;
    return 0;
}

static int seg_group_check(SegGroup& G, hipStream_t st) {
    uint32_t c[4] = {0, 0, 0, 0}, e = 0;
    MPSG_CHECK(hipMemcpyAsync(c, G.d_counters, sizeof c, hipMemcpyDeviceToHost, st));
    if (G.nlarge)
        MPSG_CHECK(hipMemcpyAsync(&e, G.plan.d_pass_flags + kMaxPasses, sizeof e, hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    if (c[0]) return fail(MPSORT_ERR_LAYOUT, "a small segment exceeded the in-LDS capacity");
    if (e & 1u) return fail(MPSORT_ERR_TIMEOUT, "a look-back wait exceeded its spin limit");
    if (e) return fail(MPSORT_ERR_LAYOUT, "segmented sort: device flag 0x%x", e);
    return 0;
}

// pick the tile shape and the in-LDS threshold by timing the candidates on
// this group (cached per record size, key width and size class)
static int seg_group_tune(SegGroup& G, DevCtx& c) {
    seg_tune_load();
    const std::string key = seg_tune_key(G);
    auto it = seg_tune_cache().find(key);
    if (it != seg_tune_cache().end()) {
        G.tune = it->second;
        G.tuned = true;
        return seg_group_prepare(G, c.st);
    }
    const int hi = G.rw == 1 ? 32 : (G.rw == 2 ? 32 : (G.rw == 4 ? 16 : 8));
    const TileCfg tiles[] = {{1024, hi / 2}, {1024, hi / 4}, {512, hi}, {256, hi}};
    const int thresholds[] = {kChunkCap, kChunkCap / 4, 0};
    float best = 1e30f;
    SegTune pick = G.tune;
    int rc = 0;
    for (const TileCfg& t : tiles) {
        if (!tile_supported(G.rw, t)) continue;
        for (int thr : thresholds) {
            G.tune = SegTune{t, thr, 0};
            if ((rc = seg_group_prepare(G, c.st))) return rc;
            float best_here = 1e30f;
            for (int rep = 0; rep < 2; rep++) {
                if ((rc = seg_group_epoch_guard(G, c.st))) return rc;
                MPSG_CHECK(hipEventRecord(c.ev[kEvStart], c.st));
                if ((rc = seg_group_enqueue(G, c.st))) return rc;
                MPSG_CHECK(hipEventRecord(c.ev[kEvEnd], c.st));
                if ((rc = seg_group_check(G, c.st))) return rc;
                best_here = std::min(best_here, (float)ev_ms(c.ev[kEvStart], c.ev[kEvEnd]));
            }
            if (best_here < best) {
                best = best_here;
                pick = G.tune;
            }
        }
    }
    // then the workgroup shape of the segments sorted in LDS (how many
    // workgroups and waves share a CU), with that tile and threshold
    const bool lds_ok = (G.rw == 2 || G.rw == 4) && G.ks.key_bytes <= 8;
    if (lds_ok && pick.threshold > 0) {
        for (int form = 1; form <= 2; form++) {
            G.tune = SegTune{pick.tile, pick.threshold, form};
            if ((rc = seg_group_prepare(G, c.st))) return rc;
            if (!G.nsmall) break;
            float best_here = 1e30f;
            for (int rep = 0; rep < 2; rep++) {
                if ((rc = seg_group_epoch_guard(G, c.st))) return rc;
                MPSG_CHECK(hipEventRecord(c.ev[kEvStart], c.st));
                if ((rc = seg_group_enqueue(G, c.st))) return rc;
                MPSG_CHECK(hipEventRecord(c.ev[kEvEnd], c.st));
                if ((rc = seg_group_check(G, c.st))) return rc;
                best_here = std::min(best_here, (float)ev_ms(c.ev[kEvStart], c.ev[kEvEnd]));
            }
            if (best_here < best) {
                best = best_here;
                pick = G.tune;
            }
        }
    }
    G.tune = pick;
    G.tuned = true;
    seg_tune_cache()[key] = pick;
    seg_tune_store();
    return seg_group_prepare(G, c.st);
}

// Host staging and the in-place guard around one device sort (run).  With
// device pointers, a packed-record call whose output overlaps its input
// gathers the caller's records from a copy: the gather reads records by
// index while it writes the output.
template <class Run>
static int staged_call(DevCtx* c, const CallShape& cs, const void* base, void* out, bool devptr, int P,
                       mpsort_gpu_stats* st, Run run) {
    int rc = 0;
    const int64_t ntot = cs.ntot;
    const void* d_in = base;
    void* d_out = out;
    const size_t bytes = (size_t)ntot * (size_t)cs.elsize;
    hipEvent_t e0 = c->ev[kEvStart], e1 = c->ev[kEvEnd];
    if (!devptr && ntot > 0) {
        if ((rc = c->din.ensure(bytes))) return rc;
        if ((rc = c->dout.ensure(bytes))) return rc;
        MPSG_CHECK(hipEventRecord(e0, c->st));
        MPSG_CHECK(hipMemcpyAsync(c->din.p, base, bytes, hipMemcpyHostToDevice, c->st));
        MPSG_CHECK(hipEventRecord(e1, c->st));
        MPSG_CHECK(hipStreamSynchronize(c->st));
        st->t_h2d_ms = ev_ms(e0, e1);
        d_in = c->din.p;
        d_out = c->dout.p;
    } else if (devptr && cs.general && ntot > 0) {
        const char *a = (const char*)base, *b = (const char*)out;
        if (a < b + bytes && b < a + bytes) {
            if ((rc = c->din.ensure(bytes))) return rc;
            MPSG_CHECK(hipMemcpyAsync(c->din.p, base, bytes, hipMemcpyDeviceToDevice, c->st));
            d_in = c->din.p;
        }
    }
    if (ntot > 0) {
        rc = run(d_in, d_out);
    } else {
        st->nvranks = P;
        c->last_P = 0;
    }
    if (!rc && !devptr && ntot > 0) {
        MPSG_CHECK(hipEventRecord(e0, c->st));
        MPSG_CHECK(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, c->st));
        MPSG_CHECK(hipEventRecord(e1, c->st));
        MPSG_CHECK(hipStreamSynchronize(c->st));
        st->t_d2h_ms = ev_ms(e0, e1);
    }
    return rc;
}

extern "C" {

void* mpiu_malloc(const char* name, size_t size, const char* file, const int line) {
    return g_hmalloc(name, size, file, line, g_hud);
}
void mpiu_free(void* ptr, const char* file, const int line) {
    if (ptr) g_hfree(ptr, file, line, g_hud);
}
void mpiu_set_malloc(mpiu_malloc_func m, mpiu_free_func f, void* userdata) {
    g_hmalloc = m ? m : host_malloc_default;
    g_hfree = f ? f : host_free_default;
    g_hud = userdata;
}

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
    rc = shape_for(key, elsize, flags, devptr ? base : nullptr, devptr ? out : nullptr, ntot, &cs);
    if (rc) return rc;
    if (!cs.general && cs.rw == 8 && cs.ks.key_bytes <= 8 && !cs.exchange && cs.first_mode != 1 &&
        ntot >= kFirstMinAvg * P) {
        // 32-byte records of large ranks: sorted as packed {key, index}
        // records, so that the bucket FirstSort and the merged SecondSort
        // take them (a quarter of the radix passes' traffic, gather included)
        if ((rc = shape_for(key, elsize, flags | MPSORT_GPU_FORCE_GENERAL, devptr ? base : nullptr,
                            devptr ? out : nullptr, ntot, &cs)))
            return rc;
    }
    cs.P = P;
    cs.ntot = ntot;
    if (ntot > 0 && (!base || !out)) return fail(MPSORT_ERR_ARG, "NULL data pointer");
    DevCtx* c = nullptr;
    if ((rc = get_ctx(&c))) return rc;
    mpsort_gpu_stats st;
    memset(&st, 0, sizeof st);
    remember_key(*c, key, 0, 0);
    rc = staged_call(c, cs, base, out, devptr, P, &st, [&](const void* d_in, void* d_out) {
        HistWork w;
        int r2 = histwork_cached(*c, cs, histsort_need(cs, nin, nout, true), &w);
        return r2 ? r2 : run_histsort(*c, cs, d_in, d_out, nin, nout, nullptr, &st, w);
    });
    g_last = st;
    return rc;
}

int mpsort_gpu_newarray_noalloc(const void* base, const int64_t* nin, int P, void* out, const int64_t* nout,
                                size_t elsize, const mpsort_key_desc* key, unsigned flags) {
    g_err[0] = 0;
    int64_t ntot = 0;
    int rc = validate_counts(nin, P, nout, &ntot);
    if (rc) return rc;
    if (flags & (MPSORT_GPU_LEARNED | MPSORT_GPU_LEARNED_EXCHANGE))
        return fail(MPSORT_ERR_UNSUPPORTED, "the allocation-free entry point runs the plain path only");
    flags |= MPSORT_GPU_DEVICE_PTRS;
    CallShape cs;
    memset(&cs, 0, sizeof cs);
    if ((rc = shape_for(key, elsize, flags, base, out, ntot, &cs))) return rc;
    cs.P = P;
    cs.ntot = ntot;
    if (ntot > 0 && (!base || !out)) return fail(MPSORT_ERR_ARG, "NULL data pointer");
    DevCtx* c = nullptr;
    if ((rc = get_ctx(&c))) return rc;
    if (!c->scratch)
        return fail(MPSORT_ERR_HIP, "the scratch array of device %d has no address (HIP error %d)", c->device,
                    (int)c->scratch_err);
    mpsort_gpu_stats st;
    memset(&st, 0, sizeof st);
    remember_key(*c, key, 0, 0);
    if (ntot == 0) {
        st.nvranks = P;
        c->last_P = 0;
        g_last = st;
        return 0;
    }
    // the workspaces, carved from the scratch array: tables, look-back
    // flags, sort-record buffers, and (in-place packed records) a copy of
    // the input for the final gather
    const HistNeed need = histsort_need(cs, nin, nout);
    const size_t bytes = (size_t)ntot * elsize;
    const char *a = (const char*)base, *b = (const char*)out;
    const size_t incopy = (cs.general && a < b + bytes && b < a + bytes) ? bytes : 0;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t total = up(need.small) + up(need.lookback) + up(need.s0) + up(need.s1) + up(need.rbuf) + up(incopy);
    if (total > kScratchBytes)
        return fail(MPSORT_ERR_NOMEM, "this sort needs %zu bytes of workspace; the allocation-free entry point has "
                                      "%zu (MPSORT_GPU_NOALLOC_SCRATCH_BYTES)", total, kScratchBytes);
    static int64_t host_tab[7 * 4096 + 4];
    char* p = (char*)c->scratch;
    HistWork w;
    w.small = p;
    p += up(need.small);
    w.lookback = (uint64_t*)p;
    p += up(need.lookback);
    w.s0 = need.s0 ? (void*)p : nullptr;
    p += up(need.s0);
    w.s1 = p;
    p += up(need.s1);
    w.rbuf = need.rbuf ? (void*)p : nullptr;
    p += up(need.rbuf);
    w.host_tab = host_tab;
    w.clear_lookback = true;
    const void* d_in = base;
    if (incopy) {
        MPSG_CHECK(hipMemcpyAsync(p, base, bytes, hipMemcpyDeviceToDevice, c->st));
        d_in = p;
    }
    rc = run_histsort(*c, cs, d_in, out, nin, nout, nullptr, &st, w);
    g_last = st;
    return rc;
}

// A radix longer than the packed sort record takes (rsize above 28 bytes,
// 24 with a 64-bit index): every radix is materialised once by host
// threads, then stable passes over its 24-byte chunks, least significant
// first (an LSD sort whose digits are whole chunks: radixsort.c:146-195
// compares the radix as one little-endian integer), each sorting 32-byte
// records {chunk, original index} in place through mpsort_gpu_newarray --
// one rank for the lower chunks, the caller's ranks for the top one, whose
// counts and splitters are the ones recorded.  The caller's records are
// gathered by index at the end; both checksums are the host's.
static int sort_long_radix(const void* base, const int64_t* nin, int P, void* out, const int64_t* nout,
                           size_t elsize, void (*radix)(const void* ptr, void* radix, void* arg), size_t rsize,
                           void* arg, unsigned flags, int64_t ntot) {
    constexpr size_t kChunk = 24;
    struct Rec {
        uint64_t k[3];  // the chunk, little-endian words (the last most significant)
        uint64_t idx;   // the record's input position
    };
    static_assert(sizeof(Rec) == 32, "chunk record");
    const int nch = (int)((rsize + kChunk - 1) / kChunk);
    if (ntot == 0) {
        memset(&g_last, 0, sizeof g_last);
        return 0;
    }
    unsigned char* rad = (unsigned char*)mpiu_malloc("mpsort_gpu:radix", (size_t)ntot * rsize, __FILE__, __LINE__);
    Rec* recs = (Rec*)mpiu_malloc("mpsort_gpu:chunks", (size_t)ntot * sizeof(Rec), __FILE__, __LINE__);
    unsigned char* src = nullptr;  // the input, copied when the output overlaps it
    const size_t bytes = (size_t)ntot * elsize;
    const bool overlap = (const unsigned char*)out < (const unsigned char*)base + bytes &&
                         (const unsigned char*)base < (const unsigned char*)out + bytes;
    if (overlap) src = (unsigned char*)mpiu_malloc("mpsort_gpu:input", bytes, __FILE__, __LINE__);
    auto release = [&]() {
        if (rad) mpiu_free(rad, __FILE__, __LINE__);
        if (recs) mpiu_free(recs, __FILE__, __LINE__);
        if (src) mpiu_free(src, __FILE__, __LINE__);
    };
    if (!rad || !recs || (overlap && !src)) {
        release();
        return fail(MPSORT_ERR_NOMEM, "host allocation for a %zu-byte radix failed", rsize);
    }
    const int T = host_threads(ntot);
    std::vector<long long> part(T, 0);
    parallel_ranges(ntot, T, [&](int t, int64_t lo, int64_t hi) {
        long long sum = 0;
        for (int64_t i = lo; i < hi; i++) {
            const unsigned char* rec = (const unsigned char*)base + (size_t)i * elsize;
            memset(rad + (size_t)i * rsize, 0, rsize);
            radix(rec, rad + (size_t)i * rsize, arg);
            recs[i].idx = (uint64_t)i;
            for (size_t b = 0; b < elsize; b++) sum += (signed char)rec[b];
        }
        part[t] = sum;
    });
    long long sum_in = 0;
    for (long long v : part) sum_in += v;
    if (src) memcpy(src, base, bytes);
    const unsigned char* from = src ? src : (const unsigned char*)base;
    mpsort_key_desc key;
    memset(&key, 0, sizeof key);
    key.offset = 0;
    key.width = 8;
    key.nwords = 3;
    const unsigned pass_flags = flags & (MPSORT_GPU_NO_CHECKSUM | MPSORT_GPU_TIMING);
    int rc = 0;
    for (int ch = 0; ch < nch && !rc; ch++) {
        const size_t off = (size_t)ch * kChunk, len = std::min(kChunk, rsize - off);
        parallel_ranges(ntot, T, [&](int, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; i++) {
                Rec& r = recs[i];
                r.k[0] = r.k[1] = r.k[2] = 0;
                memcpy(r.k, rad + (size_t)r.idx * rsize + off, len);
            }
        });
        const bool top = ch == nch - 1;
        const int64_t one = ntot;
        rc = mpsort_gpu_newarray(recs, top ? nin : &one, top ? P : 1, recs, top ? nout : &one, sizeof(Rec), &key,
                                 pass_flags);
    }
    if (!rc) {
        parallel_ranges(ntot, T, [&](int t, int64_t lo, int64_t hi) {
            long long sum = 0;
            for (int64_t i = lo; i < hi; i++) {
                unsigned char* dst = (unsigned char*)out + (size_t)i * elsize;
                memcpy(dst, from + (size_t)recs[i].idx * elsize, elsize);
                for (size_t b = 0; b < elsize; b++) sum += (signed char)dst[b];
            }
            part[t] = sum;
        });
        long long sum_out = 0;
        for (long long v : part) sum_out += v;
        g_last.checksum_in = sum_in;
        g_last.checksum_out = sum_out;
        if (!(flags & MPSORT_GPU_NO_CHECKSUM) && sum_in != sum_out)
            rc = fail(MPSORT_ERR_CHECKSUM, "MPSort: Data changed after sorting; checksum mismatch (%lld != %lld)",
                      sum_in, sum_out);
    }
    release();
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
    if (elsize == 0) return fail(MPSORT_ERR_ARG, "elsize is 0");
    if (flags & MPSORT_GPU_DEVICE_PTRS) return fail(MPSORT_ERR_ARG, "the radix-callback form takes host pointers");
    const int iwords = ntot >= ((int64_t)1 << 32) ? 2 : 1;
    if (rsize == 0) return fail(MPSORT_ERR_UNSUPPORTED, "rsize is 0");
    const int rw = rsize <= 32 ? packed_words((int64_t)rsize, iwords) : 0;
    if (ntot > 0 && (!base || !out)) return fail(MPSORT_ERR_ARG, "NULL data pointer");
    if (!rw) return sort_long_radix(base, nin, P, out, nout, elsize, radix, rsize, arg, flags, ntot);
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
    cs.rw = rw;
    cs.iwords = iwords;
    fill_keyspec(&cs.ks, cs.rw, 0, (int)rsize, 0, 0);
    cs.timing = (flags & MPSORT_GPU_TIMING) != 0;
    remember_key(*c, nullptr, 1, (int64_t)rsize);
    if (ntot == 0) {
        g_last = st;
        c->last_P = 0;
        return 0;
    }
    // the radix of every record, materialised once, with its position; and
    // the input checksum -- both by host threads over record ranges
    const size_t rb = (size_t)rw * 4;
    uint32_t* packed = (uint32_t*)mpiu_malloc("mpsort_gpu:radix", (size_t)ntot * rb, __FILE__, __LINE__);
    if (!packed) return fail(MPSORT_ERR_NOMEM, "host allocation of %zu bytes failed", (size_t)ntot * rb);
    const int T = host_threads(ntot);
    std::vector<long long> part(T, 0);
    parallel_ranges(ntot, T, [&](int t, int64_t lo, int64_t hi) {
        unsigned char tmp[64];
        long long sum = 0;
        for (int64_t i = lo; i < hi; i++) {
            const unsigned char* rec = (const unsigned char*)base + (size_t)i * elsize;
            memset(tmp, 0, sizeof tmp);
            radix(rec, tmp, arg);
            uint32_t* w = packed + (size_t)i * rw;
            memcpy(w, tmp, rb - 4 * (size_t)iwords);
            if (iwords == 2) {
                w[rw - 2] = (uint32_t)i;
                w[rw - 1] = (uint32_t)((uint64_t)i >> 32);
            } else {
                w[rw - 1] = (uint32_t)i;
            }
            for (size_t b = 0; b < elsize; b++) sum += (signed char)rec[b];
        }
        part[t] = sum;
    });
    long long sum_in = 0;
    for (long long v : part) sum_in += v;
    const size_t bytes = (size_t)ntot * elsize;
    if ((rc = c->din.ensure(bytes)) || (rc = c->dout.ensure(bytes))) {
        mpiu_free(packed, __FILE__, __LINE__);
        return rc;
    }
    if (hipMemcpy(c->din.p, base, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        mpiu_free(packed, __FILE__, __LINE__);
        return fail(MPSORT_ERR_HIP, "copy of the records to the device failed");
    }
    // the gather's device checksum covers the output; the input sum is the
    // host's (the device sees only the packed records)
    cs.checksum = !(flags & MPSORT_GPU_NO_CHECKSUM);
    HistWork w;
    rc = histwork_cached(*c, cs, histsort_need(cs, nin, nout, true), &w);
    if (!rc) rc = run_histsort(*c, cs, c->din.p, c->dout.p, nin, nout, packed, &st, w);
    mpiu_free(packed, __FILE__, __LINE__);
    if (!rc) {
        MPSG_CHECK(hipMemcpy(out, c->dout.p, bytes, hipMemcpyDeviceToHost));
        st.checksum_in = sum_in;
        if (cs.checksum && sum_in != st.checksum_out)
            rc = fail(MPSORT_ERR_CHECKSUM, "MPSort: Data changed after sorting; checksum mismatch (%lld != %lld)",
                      sum_in, st.checksum_out);
    }
    g_last = st;
    return rc;
}

int mpsort_gpu_last_stats(mpsort_gpu_stats* s) {
    if (!s) return MPSORT_ERR_ARG;
    *s = g_last;
    return 0;
}

// The reference's timer lines in its order (mpsort-mpi.c:366-602): seconds
// since the previous line.  Stages that one kernel does together with the
// next stage print 0: PmaxPmin runs inside the bisection kernel (its time is
// under findP, which covers every bisection round; bisectNNNN names the
// round count), LayDistr inside the layout solve.
void mpsort_gpu_report_last_run(void) {
    const mpsort_gpu_stats& s = g_last;
    printf("FirstSort: %g\n", s.t_first_sort_ms * 1e-3);
    printf("PmaxPmin: %g\n", 0.0);
    printf("bisect%04d: %g\n", s.bisect_rounds, 0.0);
    printf("findP: %g\n", s.t_findp_ms * 1e-3);
    printf("LayDistr: %g\n", 0.0);
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
    c.s2.release();
    c.lsmall.release();
    c.ovf.release();
    c.small.release();
    c.lookback.release();
    c.ltile.release();
    c.mbnd.release();
    c.mdesc.release();
    c.fsamp.release();
    c.fsplan.release();
    c.fsplan_key = 0;
    c.hmirror.release();
    c.gstage.release();
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
int mpsort_gpu_set_faults(unsigned faults, int64_t bucket_cap) {
    const unsigned known = MPSORT_GPU_FAULT_SPLITTERS | MPSORT_GPU_FAULT_BUCKET_CAP | MPSORT_GPU_FAULT_PLAIN |
                           MPSORT_GPU_FAULT_CORRUPT;
    if (faults & ~known) return fail(MPSORT_ERR_ARG, "unknown fault bits 0x%x", faults & ~known);
    if ((faults & MPSORT_GPU_FAULT_BUCKET_CAP) && (bucket_cap < 1 || bucket_cap > kChunkCap))
        return fail(MPSORT_ERR_ARG, "bucket_cap %lld outside [1, %d]", (long long)bucket_cap, kChunkCap);
    g_faults = faults;
    g_fault_cap = (faults & MPSORT_GPU_FAULT_BUCKET_CAP) ? bucket_cap : 0;
    return MPSORT_OK;
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


// ---- learned-splitter path --------------------------------------------------
struct mpsort_gpu_model_s {
    LearnedModel m;
};

static uint64_t hash64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// deterministic initial parameters: h1_j = tanh(8 (x - c_j)) with centres c_j
// spread over [0, 1]; small positive W2, spread b2; w3 ~ 1/32; b3 = 0.5
// Initial root: hidden unit j of the first layer is a soft step centred on
// the sample quantile (j + 1/2) / 32 of x, as steep as the gap between the
// neighbouring quantiles (xs: the S training inputs, sorted); the second
// layer starts as a small positive mixing.
static void mlp_init(float* raw, uint64_t seed, const float* xs, int64_t S) {
    auto inv_softplus = [](double v) { return (float)std::log(std::expm1(v)); };
    auto xq = [&](double f) {
        int64_t i = (int64_t)(f * (double)S);
        i = i < 0 ? 0 : (i >= S ? S - 1 : i);
        return (double)xs[i];
    };
    for (int j = 0; j < kH; j++) {
        const double c = xq((j + 0.5) / kH);
        const double gap = std::max(xq((j + 1.0) / kH) - xq((double)j / kH), 1e-6);
        const double w = std::min(std::max(2.0 / gap, 1.0), 1e4);
        raw[kW1 + j] = inv_softplus(w);
        raw[kB1 + j] = (float)(-w * c);
    }
    for (int q = 0; q < kH * kH; q++) {
        const double u = (double)(hash64(seed + 7919 * (uint64_t)q) >> 11) * 0x1.0p-53;
        raw[kW2 + q] = inv_softplus(0.05 + 0.1 * u);
    }
    for (int k = 0; k < kH; k++) {
        const double u = (double)(hash64(seed + 104729 + (uint64_t)k) >> 11) * 0x1.0p-53;
        raw[kB2 + k] = (float)(-1.6 + 3.2 * u);
        raw[kW3 + k] = inv_softplus(1.0 / 64.0);
    }
    raw[kB3] = 0.5f;
}

int mpsort_gpu_model_train(mpsort_gpu_model** out, const void* sorted, int64_t n, size_t elsize,
                           const mpsort_key_desc* key, int64_t nsample, uint64_t seed, unsigned flags) {
    g_err[0] = 0;
    if (!out) return fail(MPSORT_ERR_ARG, "model output pointer is NULL");
    *out = nullptr;
    if (!key || !sorted || n < 2 || elsize == 0) return fail(MPSORT_ERR_ARG, "bad training arguments");
    if (key->width != 4 && key->width != 8) return fail(MPSORT_ERR_UNSUPPORTED, "key width %d", key->width);
    const int64_t kb = (int64_t)key->width * key->nwords;
    if (kb > 8) return fail(MPSORT_ERR_UNSUPPORTED, "the learned path handles keys of at most 8 bytes");
    if (key->offset < 0 || key->offset + kb > (int64_t)elsize) return fail(MPSORT_ERR_ARG, "key outside record");
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = c->st;
    const bool devptr = (flags & MPSORT_GPU_DEVICE_PTRS) != 0;
    const size_t bytes = (size_t)n * elsize;
    const void* d_sorted = sorted;
    void* tmpbuf = nullptr;
    if (!devptr) {
        tmpbuf = dev_alloc(bytes);
        if (!tmpbuf) return fail(MPSORT_ERR_NOMEM, "device allocation failed");
        MPSG_CHECK(hipMemcpy(tmpbuf, sorted, bytes, hipMemcpyHostToDevice));
        d_sorted = tmpbuf;
    }
    uint64_t xm = 0;
    if (key->is_signed)
        for (int b = key->width - 1; b < kb; b += key->width) xm |= 0x80ull << (8 * b);
    auto key_at = [&](int64_t pos, uint64_t* k) -> int {
        unsigned char buf[8] = {0};
        MPSG_CHECK(hipMemcpy(buf, (const char*)d_sorted + pos * elsize + key->offset, (size_t)kb,
                             hipMemcpyDeviceToHost));
        uint64_t v = 0;
        for (int b = 0; b < kb; b++) v |= (uint64_t)buf[b] << (8 * b);
        *k = v ^ xm;
        return 0;
    };
    mpsort_gpu_model* mo = new mpsort_gpu_model_s();
    LearnedModel& M = mo->m;
    M.device = c->device;
    M.key_bytes = (int)kb;
    if ((rc = key_at(0, &M.kmin)) || (rc = key_at(n - 1, &M.kmax))) { delete mo; dev_free(tmpbuf); return rc; }
    if (M.kmax < M.kmin) std::swap(M.kmin, M.kmax);
    M.inv_range = M.kmax > M.kmin ? 1.0 / (double)(M.kmax - M.kmin) : 0.0;
    M.log_scale = M.kmax > M.kmin ? 1.0 / std::log2((double)(M.kmax - M.kmin) + 1.0) : 0.0;
    const int64_t S = std::max<int64_t>(16, std::min<int64_t>(nsample > 0 ? nsample : (1 << 20), n));
    M.nsample = S;
    M.L = (int)std::max<int64_t>(1, std::min<int64_t>(32768, S / 8));
    M.linear_root = (flags & MPSORT_GPU_MODEL_LINEAR_ROOT) ? 1 : 0;

    const int grad_blocks = mlp_grad_blocks(S, 1, 0);
    M.raw = (float*)dev_alloc(sizeof(float) * kMlpParams);
    M.eff = (float*)dev_alloc(sizeof(float) * kMlpParams);
    M.leaves = (float*)dev_alloc(sizeof(float) * 4 * (size_t)M.L);
    M.xstart = (float*)dev_alloc(sizeof(float) * ((size_t)M.L + 1));
    float* am = (float*)dev_alloc(sizeof(float) * kMlpParams);
    float* av = (float*)dev_alloc(sizeof(float) * kMlpParams);
    float* partial = (float*)dev_alloc(sizeof(float) * kMlpParams * (size_t)grad_blocks);
    float* lossp = (float*)dev_alloc(sizeof(float) * (size_t)grad_blocks);
    uint64_t* skeys = (uint64_t*)dev_alloc(sizeof(uint64_t) * S);
    float* x = (float*)dev_alloc(sizeof(float) * S);
    float* y = (float*)dev_alloc(sizeof(float) * S);
    int* leaf = (int*)dev_alloc(sizeof(int) * S);
    int64_t* start = (int64_t*)dev_alloc(sizeof(int64_t) * (M.L + 1));
    float* fval = (float*)dev_alloc(sizeof(float) * S);
    auto cleanup = [&]() {
        dev_free(am); dev_free(av); dev_free(partial); dev_free(lossp); dev_free(skeys); dev_free(x); dev_free(y);
        dev_free(leaf); dev_free(start); dev_free(fval); dev_free(tmpbuf);
    };
    if (!M.raw || !M.eff || !M.leaves || !M.xstart || !am || !av || !partial || !lossp || !skeys || !x || !y || !leaf || !start ||
        !fval) {
        cleanup();
        mpsort_gpu_model_free(mo);
        return fail(MPSORT_ERR_NOMEM, "device allocation failed (model)");
    }
    auto run = [&]() -> int {
        int r2 = launch_model_sample(d_sorted, (int64_t)elsize, key->offset, (int)kb, xm, n, S, seed, M.dev(),
                                     skeys, x, y, st);
        if (r2) return r2;
        std::vector<uint64_t> hk(S);
        MPSG_CHECK(hipMemcpyAsync(hk.data(), skeys, sizeof(uint64_t) * S, hipMemcpyDeviceToHost, st));
        MPSG_CHECK(hipStreamSynchronize(st));
        // root input: linear or log-scaled key, whichever spreads the samples
        // more evenly over L equal cells of x (fewer samples in the fullest)
        {
            std::vector<int32_t> cl(M.L), cg(M.L);
            for (int64_t i = 0; i < S; i++) {
                const double d = (double)(hk[i] - M.kmin);
                const double xl = std::min(1.0, d * M.inv_range), xg = std::min(1.0, std::log2(d + 1.0) * M.log_scale);
                cl[std::min<int64_t>(M.L - 1, (int64_t)(xl * M.L))]++;
                cg[std::min<int64_t>(M.L - 1, (int64_t)(xg * M.L))]++;
            }
            const int32_t ml = *std::max_element(cl.begin(), cl.end()), mg = *std::max_element(cg.begin(), cg.end());
            if (mg < ml) {
                M.log_input = 1;
                if ((r2 = launch_model_sample(d_sorted, (int64_t)elsize, key->offset, (int)kb, xm, n, S, seed,
                                              M.dev(), skeys, x, y, st)))
                    return r2;
            }
        }
        // heavy keys: runs of at least max(4, S / 2^18) equal sampled keys,
        // the kMaxHeavy longest, each with its quantile range [i / S, j / S]
        // (samples i .. j - 1 of the stratified sample)
        const int64_t hthr = std::max<int64_t>(4, S >> 18);
        struct Run {
            int64_t i, j;
            uint64_t v;
        };
        std::vector<Run> runs;
        for (int64_t i = 0; i < S;) {
            int64_t j = i;
            while (j < S && hk[j] == hk[i]) j++;
            if (j - i >= hthr) runs.push_back({i, j, hk[i]});
            i = j;
        }
        if (runs.size() > (size_t)kMaxHeavy) {
            std::sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.j - a.i > b.j - b.i; });
            runs.resize(kMaxHeavy);
            std::sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.v < b.v; });
        }
        std::vector<uint64_t> hv;
        std::vector<float2> hq;
        for (auto& r : runs) {
            hv.push_back(r.v);
            hq.push_back(make_float2((float)((double)r.i / (double)S), (float)((double)r.j / (double)S)));
        }
        M.nheavy = (int)hv.size();
        M.heavy = (uint64_t*)dev_alloc(sizeof(uint64_t) * (hv.size() + 1));
        M.heavy_q = (float2*)dev_alloc(sizeof(float2) * (hv.size() + 1));
        if (!M.heavy || !M.heavy_q) return fail(MPSORT_ERR_NOMEM, "device allocation failed (heavy keys)");
        if (!hv.empty()) {
            MPSG_CHECK(hipMemcpy(M.heavy, hv.data(), sizeof(uint64_t) * hv.size(), hipMemcpyHostToDevice));
            MPSG_CHECK(hipMemcpy(M.heavy_q, hq.data(), sizeof(float2) * hq.size(), hipMemcpyHostToDevice));
        }
        // the root network
        float init[kMlpParams];
        {
            std::vector<float> hx(S);
            MPSG_CHECK(hipMemcpyAsync(hx.data(), x, sizeof(float) * S, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipStreamSynchronize(st));
            mlp_init(init, seed, hx.data(), S);
        }
        MPSG_CHECK(hipMemcpy(M.raw, init, sizeof init, hipMemcpyHostToDevice));
        MPSG_CHECK(hipMemset(am, 0, sizeof(float) * kMlpParams));
        MPSG_CHECK(hipMemset(av, 0, sizeof(float) * kMlpParams));
        if (!M.linear_root) {
            // Adam, 2000 steps over 16 interleaved mini-batches, learning rate
            // cosine-annealed from 1e-2 to 1e-4
            const int stride = 16, steps = 2000;
            for (int step = 1; step <= steps; step++) {
                const int off = (step - 1) % stride;
                const float lr = 1e-4f + 0.5f * (1e-2f - 1e-4f) * (1.f + std::cos(3.14159265f * step / steps));
                if ((r2 = launch_mlp_effective(M.raw, M.eff, st))) return r2;
                if ((r2 = launch_mlp_grad(x, y, S, stride, off, M.raw, M.eff, partial, lossp, st))) return r2;
                if ((r2 = launch_mlp_adam(M.raw, partial, mlp_grad_blocks(S, stride, off), am, av, lr, step, nullptr,
                                          st)))
                    return r2;
            }
            M.steps = steps;
            // loss of the final root over every sample
            if ((r2 = launch_mlp_effective(M.raw, M.eff, st))) return r2;
            if ((r2 = launch_mlp_grad(x, y, S, 1, 0, M.raw, M.eff, partial, lossp, st))) return r2;
            std::vector<float> lp(grad_blocks);
            MPSG_CHECK(hipMemcpyAsync(lp.data(), lossp, sizeof(float) * grad_blocks, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipStreamSynchronize(st));
            double L2 = 0;
            for (float v : lp) L2 += v;
            M.loss = (float)L2;
        } else {
            // the linear root's mean squared error (r = x)
            std::vector<float> hx(S), hy(S);
            MPSG_CHECK(hipMemcpyAsync(hx.data(), x, sizeof(float) * S, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipMemcpyAsync(hy.data(), y, sizeof(float) * S, hipMemcpyDeviceToHost, st));
            MPSG_CHECK(hipStreamSynchronize(st));
            double L2 = 0;
            for (int64_t i = 0; i < S; i++) L2 += ((double)hx[i] - hy[i]) * ((double)hx[i] - hy[i]);
            M.loss = (float)(L2 / (double)S);
        }
        if ((r2 = launch_mlp_effective(M.raw, M.eff, st))) return r2;
        if ((r2 = launch_leaves(x, y, S, M.eff, M.L, M.linear_root, leaf, start, M.leaves, st))) return r2;
        if ((r2 = launch_leaf_xstart(M.eff, M.L, M.linear_root, M.xstart, st))) return r2;
        std::vector<int64_t> hs(M.L + 1);
        MPSG_CHECK(hipMemcpyAsync(hs.data(), start, sizeof(int64_t) * (M.L + 1), hipMemcpyDeviceToHost, st));
        // error bound over the samples
        if ((r2 = launch_model_eval(M.dev(), skeys, S, fval, st))) return r2;
        std::vector<float> fv(S), yv(S);
        MPSG_CHECK(hipMemcpyAsync(fv.data(), fval, sizeof(float) * S, hipMemcpyDeviceToHost, st));
        MPSG_CHECK(hipMemcpyAsync(yv.data(), y, sizeof(float) * S, hipMemcpyDeviceToHost, st));
        MPSG_CHECK(hipStreamSynchronize(st));
        float e = 0.f;
        for (int64_t i = 0; i < S; i++) e = std::max(e, std::fabs(fv[i] - yv[i]));
        M.err_max = e;
        int64_t ml = 0;
        for (int l = 0; l < M.L; l++) ml = std::max<int64_t>(ml, hs[l + 1] - hs[l]);
        M.max_leaf = (int)ml;
        return 0;
    };
    rc = run();
    cleanup();
    if (rc) {
        mpsort_gpu_model_free(mo);
        return rc;
    }
    M.train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = mo;
    return 0;
}

int mpsort_gpu_model_free(mpsort_gpu_model* mo) {
    if (!mo) return 0;
    dev_free(mo->m.raw);
    dev_free(mo->m.eff);
    dev_free(mo->m.leaves);
    dev_free(mo->m.xstart);
    dev_free(mo->m.heavy);
    dev_free(mo->m.heavy_q);
    delete mo;
    return 0;
}

int mpsort_gpu_model_get_info(const mpsort_gpu_model* mo, mpsort_gpu_model_info* info) {
    if (!mo || !info) return fail(MPSORT_ERR_ARG, "NULL argument");
    const LearnedModel& M = mo->m;
    memset(info, 0, sizeof *info);
    info->nsample = M.nsample;
    info->leaves = M.L;
    info->linear_root = M.linear_root;
    info->steps = M.steps;
    info->nheavy = M.nheavy;
    info->max_leaf_samples = M.max_leaf;
    info->log_input = M.log_input;
    info->train_ms = M.train_ms;
    info->loss = M.loss;
    info->err_max = M.err_max;
    info->kmin = M.kmin;
    info->kmax = M.kmax;
    return 0;
}

int mpsort_gpu_model_eval(const mpsort_gpu_model* mo, const uint64_t* keys, int64_t n, float* out) {
    if (!mo || (n > 0 && (!keys || !out))) return fail(MPSORT_ERR_ARG, "NULL argument");
    if (n <= 0) return 0;
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    uint64_t* dk = (uint64_t*)dev_alloc(sizeof(uint64_t) * n);
    float* dv = (float*)dev_alloc(sizeof(float) * n);
    if (!dk || !dv) { dev_free(dk); dev_free(dv); return fail(MPSORT_ERR_NOMEM, "device allocation failed"); }
    rc = 0;
    if (hipMemcpy(dk, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice) != hipSuccess) rc = MPSORT_ERR_HIP;
    if (!rc) rc = launch_model_eval(mo->m.dev(), dk, n, dv, c->st);
    if (!rc && hipStreamSynchronize(c->st) != hipSuccess) rc = MPSORT_ERR_HIP;
    if (!rc && hipMemcpy(out, dv, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = MPSORT_ERR_HIP;
    dev_free(dk);
    dev_free(dv);
    return rc ? fail(rc, "model evaluation failed") : 0;
}

int mpsort_gpu_model_params(const mpsort_gpu_model* mo, float* raw, float* leaves) {
    if (!mo) return fail(MPSORT_ERR_ARG, "NULL model");
    if (raw) MPSG_CHECK(hipMemcpy(raw, mo->m.raw, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (leaves) MPSG_CHECK(hipMemcpy(leaves, mo->m.leaves, sizeof(float) * 4 * (size_t)mo->m.L, hipMemcpyDeviceToHost));
    return 0;
}

int mpsort_gpu_mlp_grad(const float* raw, const float* x, const float* y, int64_t n, float* grad, float* loss) {
    if (!raw || !x || !y || !grad || n <= 0) return fail(MPSORT_ERR_ARG, "bad arguments");
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    const int blocks = mlp_grad_blocks(n, 1, 0);
    float* d_raw = (float*)dev_alloc(sizeof(float) * kMlpParams);
    float* d_eff = (float*)dev_alloc(sizeof(float) * kMlpParams);
    float* d_x = (float*)dev_alloc(sizeof(float) * n);
    float* d_y = (float*)dev_alloc(sizeof(float) * n);
    float* d_part = (float*)dev_alloc(sizeof(float) * kMlpParams * (size_t)blocks);
    float* d_loss = (float*)dev_alloc(sizeof(float) * blocks);
    float* d_g = (float*)dev_alloc(sizeof(float) * kMlpParams);
    auto done = [&](int r) {
        dev_free(d_raw); dev_free(d_eff); dev_free(d_x); dev_free(d_y); dev_free(d_part); dev_free(d_loss);
        dev_free(d_g);
        return r;
    };
    if (!d_raw || !d_eff || !d_x || !d_y || !d_part || !d_loss || !d_g) return done(fail(MPSORT_ERR_NOMEM, "alloc"));
    if (hipMemcpy(d_raw, raw, sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_x, x, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_y, y, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
        return done(fail(MPSORT_ERR_HIP, "copy failed"));
    if ((rc = launch_mlp_effective(d_raw, d_eff, c->st))) return done(rc);
    if ((rc = launch_mlp_grad(d_x, d_y, n, 1, 0, d_raw, d_eff, d_part, d_loss, c->st))) return done(rc);
    if ((rc = launch_mlp_adam(d_raw, d_part, blocks, nullptr, nullptr, 0.f, 1, d_g, c->st))) return done(rc);
    std::vector<float> lp(blocks);
    if (hipStreamSynchronize(c->st) != hipSuccess ||
        hipMemcpy(grad, d_g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(lp.data(), d_loss, sizeof(float) * blocks, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(MPSORT_ERR_HIP, "copy failed"));
    if (loss) {
        double a = 0;
        for (float v : lp) a += v;
        *loss = (float)a;
    }
    return done(0);
}

int mpsort_gpu_newarray_learned(const void* base, const int64_t* nin, int P, void* out, const int64_t* nout,
                                size_t elsize, const mpsort_key_desc* key, const mpsort_gpu_model* model,
                                unsigned flags) {
    g_err[0] = 0;
    if (!model) return fail(MPSORT_ERR_ARG, "model is NULL");
    int64_t ntot = 0;
    int rc = validate_counts(nin, P, nout, &ntot);
    if (rc) return rc;
    CallShape cs;
    memset(&cs, 0, sizeof cs);
    const bool devptr = (flags & MPSORT_GPU_DEVICE_PTRS) != 0;
    rc = shape_for(key, elsize, flags, devptr ? base : nullptr, devptr ? out : nullptr, ntot, &cs);
    if (rc) return rc;
    const int64_t kb = (int64_t)key->width * key->nwords;
    if (kb > 8 || kb != model->m.key_bytes)
        return fail(MPSORT_ERR_UNSUPPORTED, "the model was trained for %d-byte keys; this key has %lld bytes",
                    model->m.key_bytes, (long long)kb);
    if (cs.rw != 2 && cs.rw != 4) {
        // the bucket kernels take 8- and 16-byte sort records: pack {key, index}
        if ((rc = shape_for(key, elsize, flags | MPSORT_GPU_FORCE_GENERAL, nullptr, nullptr, ntot, &cs))) return rc;
    }
    cs.P = P;
    cs.ntot = ntot;
    if (ntot > 0 && (!base || !out)) return fail(MPSORT_ERR_ARG, "NULL data pointer");
    DevCtx* c = nullptr;
    if ((rc = get_ctx(&c))) return rc;
    mpsort_gpu_stats st;
    memset(&st, 0, sizeof st);
    remember_key(*c, key, 0, 0);
    rc = staged_call(c, cs, base, out, devptr, P, &st, [&](const void* d_in, void* d_out) {
        return run_histsort_learned(*c, cs, d_in, d_out, nin, nout, model->m, &st);
    });
    g_last = st;
    return rc;
}


// ---- batched segmented sort -------------------------------------------------
struct mpsort_gpu_segplan_s {
    int device = 0;
    unsigned flags = 0;
    std::vector<SegGroup*> groups;
    bool prepared = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int mpsort_gpu_segplan_create(mpsort_gpu_segplan** out, unsigned flags) {
    g_err[0] = 0;
    if (!out) return fail(MPSORT_ERR_ARG, "plan output pointer is NULL");
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    mpsort_gpu_segplan* p = new mpsort_gpu_segplan_s();
    p->device = c->device;
    p->flags = flags;
    *out = p;
    return 0;
}

int mpsort_gpu_segplan_add(mpsort_gpu_segplan* p, const void* d_in, void* d_out, const int64_t* seg_count, int nseg,
                           size_t elsize, const mpsort_key_desc* key) {
    g_err[0] = 0;
    if (!p || !key || (nseg > 0 && (!seg_count || !d_in || !d_out)) || nseg < 0)
        return fail(MPSORT_ERR_ARG, "NULL or negative argument");
    if (p->prepared) return fail(MPSORT_ERR_ARG, "the plan has already run; create a new plan to add segments");
    const int64_t kb = (int64_t)key->width * key->nwords;
    if (key->width != 4 && key->width != 8) return fail(MPSORT_ERR_UNSUPPORTED, "key width %d", key->width);
    if (elsize != 4 && elsize != 8 && elsize != 16 && elsize != 32)
        return fail(MPSORT_ERR_UNSUPPORTED, "batched sort: records of 4, 8, 16 or 32 bytes");
    if (key->offset < 0 || (key->offset & 3) || key->offset + kb > (int64_t)elsize || kb > kMaxKeyBytes)
        return fail(MPSORT_ERR_UNSUPPORTED, "batched sort: the key must be 4-byte aligned inside the record");
    SegGroup* G = new SegGroup();
    G->in = d_in;
    G->out = d_out;
    G->elsize = elsize;
    G->rw = (int)(elsize / 4);
    fill_keyspec(&G->ks, G->rw, (int)key->offset, (int)kb, key->width, key->is_signed);
    G->lk = make_lkey(G->ks);
    G->tune.tile = default_tile(G->rw);
    G->tune.form = (p->flags & MPSORT_GPU_SEGPLAN_SMALL_512x6) ? 2 : ((p->flags & MPSORT_GPU_SEGPLAN_SMALL_512x9) ? 1 : 0);
    G->each = (p->flags & MPSORT_GPU_SEGPLAN_EACH) != 0;
    G->nseg = nseg;
    G->count.assign(seg_count, seg_count + nseg);
    G->begin.resize(nseg);
    for (int i = 0; i < nseg; i++) {
        if (seg_count[i] < 0) {
            delete G;
            return fail(MPSORT_ERR_ARG, "negative segment size");
        }
        G->begin[i] = G->nrec;
        G->nrec += seg_count[i];
    }
    p->groups.push_back(G);
    return 0;
}

int mpsort_gpu_segplan_run(mpsort_gpu_segplan* p, mpsort_gpu_stats* stats) {
    g_err[0] = 0;
    if (!p) return fail(MPSORT_ERR_ARG, "NULL plan");
    DevCtx* c = nullptr;
    int rc = get_ctx(&c);
    if (rc) return rc;
    if (c->device != p->device) return fail(MPSORT_ERR_ARG, "the plan belongs to device %d", p->device);
    hipStream_t st = c->st;
    if (!p->prepared) {
        for (SegGroup* G : p->groups) {
            if (p->flags & MPSORT_GPU_SEGPLAN_TUNE) rc = seg_group_tune(*G, *c);
            else rc = seg_group_prepare(*G, st);
            if (rc) return rc;
        }
        p->prepared = true;
    }
    for (SegGroup* G : p->groups)
        if ((rc = seg_group_epoch_guard(*G, st))) return rc;
    MPSG_CHECK(hipEventRecord(c->ev[kEvStart], st));
    if (p->flags & MPSORT_GPU_SEGPLAN_GRAPH) {
        if (!p->exec) {
            MPSG_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (SegGroup* G : p->groups) {
                if ((rc = seg_group_enqueue(*G, st))) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
            }
            MPSG_CHECK(hipStreamEndCapture(st, &p->graph));
            MPSG_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        }
        MPSG_CHECK(hipGraphLaunch(p->exec, st));
    } else {
        for (SegGroup* G : p->groups)
            if ((rc = seg_group_enqueue(*G, st))) return rc;
    }
    MPSG_CHECK(hipEventRecord(c->ev[kEvEnd], st));
    for (SegGroup* G : p->groups)
        if ((rc = seg_group_check(*G, st))) return rc;
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->t_total_ms = ev_ms(c->ev[kEvStart], c->ev[kEvEnd]);
        int64_t n = 0;
        for (SegGroup* G : p->groups) n += G->nrec;
        stats->nrecords = n;
        stats->nvranks = 0;
    }
    return 0;
}

int mpsort_gpu_segplan_info(const mpsort_gpu_segplan* p, int group, int* nsmall, int* nlarge, int* block, int* items,
                            int* threshold) {
    if (!p || group < 0 || group >= (int)p->groups.size()) return fail(MPSORT_ERR_ARG, "no such group");
    const SegGroup* G = p->groups[group];
    if (nsmall) *nsmall = G->nsmall;
    if (nlarge) *nlarge = G->nlarge;
    if (block) *block = G->tune.tile.block;
    if (items) *items = G->tune.tile.items;
    if (threshold) *threshold = G->tune.threshold;
    return 0;
}

int mpsort_gpu_segplan_shape(const mpsort_gpu_segplan* p, int group, int* form, int* waves_per_cu) {
    if (!p || group < 0 || group >= (int)p->groups.size()) return fail(MPSORT_ERR_ARG, "no such group");
    const SegGroup* G = p->groups[group];
    if (form) *form = G->tune.form;
    if (waves_per_cu) *waves_per_cu = form_waves_per_cu(G->tune.form);
    return 0;
}

int mpsort_gpu_segplan_free(mpsort_gpu_segplan* p) {
    if (!p) return 0;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (SegGroup* G : p->groups) {
        G->small.release();
        G->plan_buf.release();
        G->tmp.release();
        G->lookback.release();
        delete G;
    }
    delete p;
    return 0;
}

int mpsort_gpu_sort_segments(const void* d_in, void* d_out, const int64_t* seg_count, int nseg, size_t elsize,
                             const mpsort_key_desc* key, unsigned flags) {
    mpsort_gpu_segplan* p = nullptr;
    int rc = mpsort_gpu_segplan_create(&p, flags);
    if (rc) return rc;
    if (!(rc = mpsort_gpu_segplan_add(p, d_in, d_out, seg_count, nseg, elsize, key))) rc = mpsort_gpu_segplan_run(p, nullptr);
    mpsort_gpu_segplan_free(p);
    return rc;
}

}  // extern "C"
