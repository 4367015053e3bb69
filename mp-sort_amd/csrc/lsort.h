// lsort.h -- learned local sort (lsort.hip) declarations.
#pragma once
#include "internal.h"

namespace mpsg {

constexpr int kLBins = 1024;    // learned buckets per segment (10-bit digit)
constexpr int kLBounds = 1023;  // boundaries per segment table -> buckets 0..1023
constexpr int kGuideBits = 12;  // bucket-search guide of the histogram pass: 4096 key cells
constexpr int kGuideCells = 1 << kGuideBits;
constexpr int kChunkCap = 9216; // largest bucket sorted in LDS (1024 threads x 9)
// merged SecondSort: buckets per receiver, table boundaries, in-LDS
// capacity of a group of buckets (three workgroups per CU) and of a single
// larger bucket (two per CU)
constexpr int kMergeBins = 3072;
constexpr int kMergeBounds = 3071;
constexpr int kMergeCap = 3072;
constexpr int kMergeBigCap = 4608;

// key of a sort record as one uint64 (keys of at most 8 bytes): bytes
// [sh/8, sh/8 + key_bytes) of words w0..w0+2, masked, XOR-ed with the sign
// masks
struct LKey {
    int32_t w0;
    int32_t sh;
    uint64_t mask;
    uint64_t xmask;
};

inline LKey make_lkey(const KeySpec& ks) {
    LKey k;
    k.w0 = ks.key_off / 4;
    k.sh = (ks.key_off % 4) * 8;
    k.mask = ks.key_bytes >= 8 ? ~0ull : ((1ull << (8 * ks.key_bytes)) - 1);
    k.xmask = 0;
    for (int b = 0; b < ks.key_bytes && b < 8; b++) k.xmask |= (uint64_t)ks.xmask[b] << (8 * b);
    return k;
}

// The merged SecondSort's bounds: bucket k of receiver r starts at
// bnd[merge_bnd_at(r, j, k, P)] in sender j's segment (k = 0 .. kMergeBins;
// one sender's bounds of one receiver side by side).
__host__ __device__ inline int64_t merge_bnd_at(int r, int j, int k, int P) {
    return ((int64_t)r * P + j) * (kMergeBins + 1) + k;
}

// Positions of the merged SecondSort's receiver-table entries in the
// senders' sorted segments, written by the FirstSort's bucket sorts while a
// sorted bucket is in LDS (keys == nullptr: not wanted).  Entry e is entry
// k - 1 of receiver r's table (e = r * kMergeBounds + k - 1); its position in
// segment j (#{keys <= entry}, relative to the segment's start) goes to
// bnd[merge_bnd_at(r, j, k, P)].
struct EntryOut {
    const uint64_t* keys;      // P * kMergeBounds entries, non-decreasing
    const uint32_t* range;     // per segment kLBins + 1: bucket b holds entries [range[b], range[b + 1])
    const int64_t* seg_begin;  // absolute first record of every segment
    int32_t* bnd;
    uint8_t* done;             // per FirstSort bucket: set once its entries are written
    int P;
};
// the entry ranges of every (segment, bucket): ftab the FirstSort tables
// (segment j's at ftab + j * ftstride), E the entries
int launch_entry_ranges(const uint64_t* ftab, int64_t ftstride, int P, const uint64_t* E, uint32_t* range,
                        hipStream_t st);
// after the FirstSort: the entries of the buckets no LDS sort handled
// (binary searches in the sorted buckets of recs) and the fixed bounds k = 0
// and k = kMergeBins
int launch_entries_fixup(int rw, const void* recs, const LKey& lk, const EntryOut& eo, const uint32_t* hist,
                         const int64_t* bucket_base, const int64_t* seg_count, hipStream_t st);

// tables: segment s uses tables + s * tstride (tstride 0: one table shared);
// tile_cnt receives every tile's 1024 bucket counts, bkt[i] the bucket of
// record i (for the scatter)
int launch_lhist(int rw, const void* src, const SegDev& sd, int64_t ntiles, int tile_recs, const LKey& lk,
                 const uint64_t* tables, int64_t tstride, uint32_t* hist, uint16_t* tile_cnt, long long* checksum,
                 uint16_t* bkt, hipStream_t st);
// nb buckets per segment: 1024 or 2048
int launch_lscan(int nseg, const uint32_t* hist, const int64_t* seg_begin, int64_t* base, hipStream_t st,
                 int nb = kLBins);
int lscatter_tile(int rw);
// tile_off[tile][b]: offset of the tile's bucket-b run inside its segment
int launch_ltile_scan(const uint16_t* tile_cnt, const SegDev& sd, const int64_t* bucket_base, uint32_t* tile_off,
                      hipStream_t st);
// records of single-value buckets go to eq_dst, the rest to dst; bkt: the
// buckets written by launch_lhist
int launch_lscatter(int rw, const void* src, void* dst, void* eq_dst, const SegDev& sd, int64_t ntiles,
                    const uint16_t* bkt, const uint64_t* tables, int64_t tstride, const uint32_t* tile_off,
                    uint32_t* err, hipStream_t st);
// Sorts every bucket (src -> dst, same slots).  eq_done: the scatter already
// wrote the single-value buckets to dst.  Buckets above kChunkCap go to the
// overflow list (*noverflow entries of {base, count}); hard (one uint32 per
// bucket) and *nhard (zeroed by the caller) list the buckets handed from the
// interpolation sort to the LSD sort.  ovf_cap (here and in the other bucket
// launchers): buckets above it take the overflow list as if they were above
// the in-LDS capacity (the capacity fault hook; kChunkCap changes nothing).
int launch_chunk_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                      const int64_t* bucket_base, const uint64_t* tables, int64_t tstride, int64_t nbuckets,
                      int64_t* overflow, uint32_t* noverflow, int max_overflow, uint32_t* hard, uint32_t* nhard,
                      long long* checksum, bool eq_done, hipStream_t st, hipEvent_t mid = nullptr,
                      uint32_t ovf_cap = (uint32_t)kChunkCap);

// SecondSort straight from the senders' sorted segments (no exchange), for
// P <= 64 receivers with kMergeBounds-boundary tables each:
//   bnd: P * P * (kMergeBins + 1) int32 (merge_bnd_at); hist / bucket_base:
//   P * kMergeBins
// positions of the table entries in the senders' whole segments (needs no
// splitters), then clamped to the receivers' windows myC with the sizes.
// With the FirstSort's tables ftab (kLBounds entries, sender j's at
// ftab + j * ftstride) and its bucket bases fbase (absolute, kLBins per
// segment), each search covers one bucket.
// win (optional, the receivers' windows myC): each search covers receiver
// r's window in the sender only; samp / spref (optional, with win): the
// receivers' samples from launch_merge_sample_tables narrow it further.
int launch_merge_bounds(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, const int64_t* seg_count,
                        int P, const uint64_t* tables, const uint64_t* ftab, int64_t ftstride, const int64_t* fbase,
                        int32_t* bnd, hipStream_t st, const int64_t* win = nullptr, const uint64_t* samp = nullptr,
                        const int64_t* spref = nullptr);
// receiver tables (P * kMergeBounds, each non-decreasing) from an even
// sample of kMergeSamples keys of every receiver's records in the senders'
// sorted segments, for callers without a model; myC: the receivers' windows
// in the senders.  samp (P * kMergeSamples) and spref (P * 65) keep the
// sample and the window prefixes for launch_merge_bounds.
constexpr int kMergeSamples = 4 * kMergeBins;
int launch_merge_sample_tables(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, int P,
                               const int64_t* myC, uint64_t* tables, uint64_t* samp, int64_t* spref,
                               hipStream_t st);
int launch_merge_count(int32_t* bnd, int P, const int64_t* myC, uint32_t* hist, hipStream_t st);
// groups of neighbouring receiver buckets (together at most kMergeCap
// records) in tasks[0 .. ntasks[0]), and the larger single buckets in
// tasks[P * kMergeBins .. + ntasks[1]) (tasks holds 2 * P * kMergeBins
// entries, ntasks zeroed before)
int launch_merge_group(const uint32_t* hist, int P, uint32_t* tasks, uint32_t* ntasks, hipStream_t st);
// a group's first bucket, its size and the table bounds of its keys
struct MergeTask {
    int64_t base;   // output position of the group
    uint32_t count; // records
    uint32_t task;  // first bucket << 3 | (buckets - 1)
    uint64_t tlo;   // keys lie in (tlo, thi]
    uint64_t thi;
};
// descriptors of the ntasks groups (ntasks = the host's copy of *ntasks):
// desc[q], and the sub-run of sender j at sub_beg[q * 64 + j] (absolute
// first record) and sub_pref[q * 64 + j] (position in the group)
int launch_merge_desc(const uint32_t* tasks, uint32_t ntasks, const int32_t* bnd, const int64_t* seg_begin, int P,
                      const int64_t* bucket_base, const uint64_t* tables, MergeTask* desc, int64_t* sub_beg,
                      int32_t* sub_pref, hipStream_t st);
// one workgroup per task of the list: big = false for the groups (up to
// kMergeCap records), true for the single larger buckets, which are sorted
// up to kMergeBigCap records; buckets above that (up to kChunkCap) are
// copied in sender order and listed in medium (*nmedium entries) for
// launch_listed_sort.
int launch_merge_bucket(int rw, bool big, const void* recs, void* dst, const LKey& lk, const int64_t* seg_begin,
                        const int32_t* bnd, int P, const uint32_t* hist, const int64_t* bucket_base,
                        const MergeTask* desc, const int64_t* sub_beg,
                        const int32_t* sub_pref, uint32_t ntasks, int64_t* overflow, uint32_t* noverflow,
                        int max_overflow, uint32_t* hard, uint32_t* nhard, uint32_t* medium, uint32_t* nmedium,
                        long long* checksum, hipStream_t st, uint32_t ovf_cap = (uint32_t)kChunkCap);
// in-LDS sort (1024-thread form) of the nlist listed buckets, src -> dst at
// the same slots (src may equal dst); nbins buckets per segment, segment s
// uses tables + s * tstride; crowded ones are appended to hard
int launch_listed_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* list, uint32_t nlist,
                       const uint32_t* hist, const int64_t* bucket_base, const uint64_t* tables, int nbins,
                       int64_t tstride, uint32_t* hard, uint32_t* nhard, long long* checksum, hipStream_t st,
                       const EntryOut* eo = nullptr);
// launch_chunk_sort with 512-thread workgroups (two per CU): buckets above
// 4096 records (up to kChunkCap) are listed in medium (*nmedium entries) for
// launch_listed_sort; no LDS radix pass is launched
int launch_chunk_sort_small(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                            const int64_t* bucket_base, const uint64_t* tables, int64_t tstride, int64_t nbuckets,
                            int64_t* overflow, uint32_t* noverflow, int max_overflow, uint32_t* hard, uint32_t* nhard,
                            uint32_t* medium, uint32_t* nmedium, long long* checksum, bool eq_done, hipStream_t st,
                            const EntryOut* eo = nullptr, uint32_t ovf_cap = (uint32_t)kChunkCap);
// independent segments (the batched sort) in LDS, one workgroup each:
// segments [0, na) of the lists cnt / base by the form chosen -- 0: 1024
// threads x 9 records (one workgroup per CU), 1: 512 x 9 (two per CU), 2:
// 512 x 6 (three per CU) -- and [na, na + nb), those above the form's
// capacity (segments_lds_cap), by the 1024-thread form; crowded ones by the
// LDS radix sort (hard: na + nb entries; counts *nhard_a, *nhard_b), which
// also takes the segments the 512 x 6 form hands on (one-value crowded bins)
int launch_segments_lds(int rw, int form, const void* src, void* dst, const LKey& lk, const uint32_t* cnt,
                        const int64_t* base, int64_t na, int64_t nb, int64_t* overflow, uint32_t* noverflow,
                        uint32_t* hard, uint32_t* nhard_a, uint32_t* nhard_b, hipStream_t st);
int segments_lds_cap(int form);
// LDS radix sort of the listed buckets (src may equal dst)
int launch_lsd_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                    const int64_t* bucket_base, const uint32_t* hard, const uint32_t* nhard, int64_t max_buckets,
                    long long* checksum, hipStream_t st);
// the listed {begin, count} segments that hold one key value: copied src ->
// dst when those differ (output checksum added); the others get the top bit
// of their count set.  max_count: the largest count in the list
int launch_one_value_filter(int rw, const void* src, void* dst, const LKey& lk, int64_t* list, uint32_t nlist,
                            int64_t max_count, long long* checksum, hipStream_t st);
constexpr int kMergeMaxPH = 64;  // largest P of the merged SecondSort

// Receiver tables made before the splitters (the model's): kEdgeReserve
// entries at each end repeat their neighbours until, with the receivers'
// windows known, launch_merge_edges spreads them evenly by value over the
// receiver's keys at or below its first and above its last other entry and
// writes their bounds
// (the model's receiver edges miss the splitters by its error; without
// them those keys would all land in the first or the last bucket)
constexpr int kEdgeReserve = 16;
int launch_tables_reserve_edges(uint64_t* tables, int P, hipStream_t st);
int launch_merge_edges(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, int P, const int64_t* myC,
                       uint64_t* tables, int32_t* bnd, hipStream_t st);

// FirstSort tables without a model: S (a multiple of kLBins) keys of every
// segment at evenly spaced positions into out (P * S; the caller sorts each
// segment's S keys), then kLBounds entries per segment taken from the
// sorted sample (sample-sort tables, same form as the model's)
int launch_seg_sample(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, const int64_t* seg_count,
                      int P, int S, uint64_t* out, hipStream_t st);
int launch_sample_tables(const uint64_t* sorted, int P, int S, uint64_t* tables, hipStream_t st);

}  // namespace mpsg
