// internal.h -- shared declarations of the MI355X histogram-sort library.
//
// Data model.  Every kernel moves "sort records": fixed-size records of
// RW 32-bit words (RW = 1, 2, 4, 8, i.e. 4/8/16/32 bytes).  When the caller's
// records already have one of these sizes and a 4-byte-aligned key they are
// sorted directly; otherwise the library packs {key bytes, u32 source index}
// sort records, sorts those and gathers the caller's records at the end.
//
// The sort key is a little-endian unsigned integer of key_bytes bytes at byte
// key_off of the record, after XOR-ing each byte with a mask (0x80 on the top
// byte of every signed word).  That integer orders records exactly as the
// reference orders their radix (radixsort.c:47-110 compares the radix as one
// little-endian unsigned integer; binding.pyx:81-121 maps signed words by
// adding 2^63, which is a top-bit flip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace mpsg {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kBins = 256;         // 8-bit digits
constexpr int kMaxKeyBytes = 32;
constexpr int kMaxPasses = kMaxKeyBytes;

// One digit pass: digit = ((rec[widx] >> shift) & 0xff) ^ xmask
struct PassDesc {
    int32_t widx;
    int32_t shift;
    uint32_t xmask;
    int32_t key_byte;  // which key byte this pass sorts on
};

struct PassTable {
    PassDesc p[kMaxPasses];
};

struct KeySpec {
    int32_t rec_words;   // RW
    int32_t key_off;     // byte offset of the key inside the sort record
    int32_t key_bytes;   // bytes of key
    int32_t reserved;
    uint8_t xmask[kMaxKeyBytes];  // per key byte XOR mask
};

// first record byte of a pass table whose npass (1 .. 8) passes are
// consecutive record bytes in order; -1 otherwise
inline int consecutive_key_bytes(const PassTable& pt, int npass) {
    if (npass < 1 || npass > 8) return -1;
    const int b0 = pt.p[0].widx * 4 + pt.p[0].shift / 8;
    for (int p = 0; p < npass; p++)
        if (pt.p[p].shift % 8 != 0 || pt.p[p].widx * 4 + pt.p[p].shift / 8 != b0 + p) return -1;
    return b0;
}

inline PassDesc make_pass(const KeySpec& ks, int key_byte) {
    PassDesc p;
    int b = ks.key_off + key_byte;
    p.widx = b / 4;
    p.shift = (b % 4) * 8;
    p.xmask = ks.xmask[key_byte];
    p.key_byte = key_byte;
    return p;
}

// Segments of records: segment s is records [begin[s], begin[s] + count[s])
// of the buffer; tiles of one segment are tile_start[s] .. tile_start[s+1]-1.
struct SegDev {
    const int64_t* begin;
    const int64_t* count;
    const int64_t* tile_start;  // nseg + 1 entries
    int32_t nseg;
    int32_t pad;
};

// Tile shape of the onesweep passes: BLOCK threads x ITEMS records each.
struct TileCfg {
    int block;
    int items;
};
bool tile_supported(int rw, TileCfg c);
// default shape for a record width; MPSORT_GPU_TILE="block:items" overrides
TileCfg default_tile(int rw);

// ---- device memory owned by the library (grow-only cache, per device) ----
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// error reporting
int set_hip_error(hipError_t e, const char* what, const char* file, int line);
#define MPSG_CHECK(call)                                                        \
    do {                                                                        \
        hipError_t _e = (call);                                                 \
        if (_e != hipSuccess) return ::mpsg::set_hip_error(_e, #call, __FILE__, __LINE__); \
    } while (0)

// ---- launchers (onesweep.hip) ----
// The exchange's runs (receiver-major, sender order): the dense table of all
// P * P (receiver, sender) slots, empty ones included, and the table of the
// non-empty runs only, with the choice between them made on the device
// (sel[0] = 1: the compact table, sel[1] = its length).  rnz / snz: per
// receiver / sender, its non-empty runs; roff: prefix of rnz.
enum ExchPolicy { kExchAuto = 0, kExchDense = 1, kExchSparse = 2 };
struct ExchRuns {
    int64_t *dst = nullptr, *src = nullptr;      // P * P
    int64_t *dst_c = nullptr, *src_c = nullptr;  // cap
    int *rnz = nullptr, *snz = nullptr, *roff = nullptr, *sel = nullptr;
    int cap = 0;
    int policy = kExchAuto;
};
struct RunTable {
    const int64_t *dst, *src, *dst_c, *src_c;
    const int* sel;
    int ndense;
};

struct SortBuffers {
    void* src;        // input sort records (read only unless == dst)
    void* dst;        // output
    void* tmp;        // ping-pong scratch, same size as the data
};

struct SegSortPlan {
    int rw;                  // record words
    TileCfg tile;
    KeySpec ks;
    int64_t nrec;            // total records over all segments
    int nseg;
    int64_t ntiles;
    // device arrays (caller owned)
    int64_t* d_begin;
    int64_t* d_count;
    int64_t* d_tile_start;
    uint32_t* d_hist;        // nseg * npass * 256
    int64_t* d_digit_base;   // nseg * npass * 256
    uint32_t* d_pass_flags;  // kMaxPasses + 1 (the last word is an error flag)
    uint64_t* d_lookback;    // ntiles * 256
    uint32_t* d_tickets;     // kMaxPasses
    long long* d_checksum;   // [0] in, [1] out
    bool sparse = false;     // segments do not tile [0, nrec): copies go per tile
    // graph-captured use: every digit pass runs (no read-back of the pass
    // flags) and pass i uses look-back epoch *d_epoch_base + i + 1
    bool all_passes = false;
    const uint32_t* d_epoch_base = nullptr;
};

inline PassTable make_pass_table(const KeySpec& ks) {
    PassTable pt;
    for (int b = 0; b < ks.key_bytes && b < kMaxPasses; b++) pt.p[b] = make_pass(ks, b);
    return pt;
}

// The two halves of segmented_sort: seg_count clears the plan's counters and
// histograms every digit place of src; seg_passes scans the histograms and
// runs the non-trivial digit passes.  A caller that builds the histograms
// itself (the fused exchange) clears them with seg_clear and calls
// seg_passes only.
int seg_clear(const SegSortPlan& plan, hipStream_t st);
int seg_count(const SegSortPlan& plan, const void* src, long long* checksum_in, hipStream_t st);
int seg_passes(const SegSortPlan& plan, const SortBuffers& bufs, hipStream_t st,
               long long* checksum_out, int* passes_run, hipEvent_t* pass_events);

// Runs the stable segmented LSD sort of `plan` over bufs.  Returns the number
// of digit passes executed in *passes_run (may be 0).  checksum_in/out select
// which signed-byte sums to accumulate into plan.d_checksum.
// pass_events, when not null, holds 2*kMaxPasses events recorded around
// every digit pass launched.
int segmented_sort(const SegSortPlan& plan, const SortBuffers& bufs, hipStream_t st,
                   bool checksum_in, bool checksum_out, int* passes_run,
                   hipEvent_t* pass_events);

int launch_copy(const void* src, void* dst, int64_t nrec, int rw, long long* checksum,
                hipStream_t st);
int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
// Many small device copies and fills in one launch (host tables staged in
// one upload, counters cleared, small results gathered for one download):
// entry i writes nwords words at dst, copied from src or, when src is
// null, equal to fill (one workgroup per entry).
constexpr int kFanMax = 24;
struct FanEntry {
    uint32_t* dst;
    const uint32_t* src;
    uint32_t nwords;
    uint32_t fill;
};
struct FanOut {
    int n;
    FanEntry e[kFanMax];
};
int launch_fanout(const FanOut& f, hipStream_t st);
int launch_add_u32(uint32_t* p, uint32_t v, hipStream_t st);  // *p += v on the device

// epoch counter for the look-back flags (process wide, per device)
uint32_t next_epoch(int device);

// workgroups of kernel f (threads each) resident at once on the current
// device; 2048 when the runtime cannot say
// (queried once per kernel, block size and device: the query costs host
// time at every launch otherwise)
template <typename F>
inline int64_t resident_blocks(F f, int threads) {
    struct Entry {
        const void* f;
        int threads, dev;
        int64_t n;
    };
    static Entry cache[64];
    static int ncache = 0;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 2048;
    for (int i = 0; i < ncache; i++)
        if (cache[i].f == (const void*)f && cache[i].threads == threads && cache[i].dev == dev) return cache[i].n;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, threads, 0) != hipSuccess || cus <= 0 || per <= 0)
        return 2048;
    const int64_t n = (int64_t)cus * per;
    if (ncache < 64) cache[ncache++] = Entry{(const void*)f, threads, dev, n};
    return n;
}

}  // namespace mpsg
