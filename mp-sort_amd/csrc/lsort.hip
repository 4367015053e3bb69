// lsort.hip -- learned local sort: the model-bucketed replacement of the
// upper radix passes of the local sorts (FirstSort mpsort-mpi.c:369,
// SecondSort :597, both radix_sort radixsort.c:35-44, a stable sort).
//
//   lhist_kernel     per-segment counts of the learned bucket of every record
//   lscatter_kernel  one stable onesweep pass whose 10-bit "digit" is the
//                    learned bucket (same ranking / look-back machinery as the
//                    radix passes in onesweep.hip, 1024 bins)
//   interp_sort_kernel one workgroup per bucket of at most kChunkCap records:
//                    records grouped into 4096 bins interpolated between the
//                    bucket's smallest and largest key, the rank of every
//                    record inside its bin, and a direct write to the
//                    bucket's final slots
//   lsd_sort_kernel  the buckets whose keys crowd one bin: stable LSD
//                    radix sort in LDS on the key bits that vary inside the
//                    bucket
// Buckets larger than kChunkCap are listed and sorted afterwards by the plain
// radix path, so the result is the stable sort whatever the model.
//
// Bucket of a key against a segment's sorted boundary table t[0..1023):
//   bucket = #{t[i] < key}, i.e. bucket b holds the keys in (t[b-1], t[b]]
// -- non-decreasing in the key.  A bucket with t[b-1] + 1 == t[b] (or b == 0
// with t[0] == 0) holds the single key value t[b]: duplicate-heavy keys get
// such a bucket of their own, and its records are final after the scatter.
#include <algorithm>

#include "device.h"
#include "lsort.h"

namespace mpsg {

// FAST: the common layout (the key in the record's first words, no bit
// shift) is taken without the word selection when the key has it (the
// condition is the same for the whole kernel; used where the registers
// allow the second code path)
template <int RW, bool FAST = false>
__device__ __forceinline__ uint64_t lkey(const uint32_t (&r)[RW], const LKey& k) {
    if (FAST && k.w0 == 0 && k.sh == 0) {
        uint64_t v = r[0];
        if constexpr (RW > 1) v |= (uint64_t)r[1] << 32;
        return (v & k.mask) ^ k.xmask;
    }
    const uint32_t a = pick_word<RW>(r, k.w0), b = pick_word<RW>(r, k.w0 + 1),
                   c = pick_word<RW>(r, k.w0 + 2);
    const uint64_t lo = (uint64_t)a | ((uint64_t)b << 32);
    const uint64_t v = k.sh ? ((lo >> k.sh) | ((uint64_t)c << (64 - k.sh))) : lo;
    return (v & k.mask) ^ k.xmask;
}

__device__ __forceinline__ uint32_t bucket_of(uint64_t key, const uint64_t* __restrict__ t) {
    int lo = 0;
#pragma unroll
    for (int step = 512; step > 0; step >>= 1)
        if (lo + step <= kLBounds && t[lo + step - 1] < key) lo += step;
    return (uint32_t)lo;
}

// bucket b of table t holds one key value only
__device__ __forceinline__ bool single_bucket(const uint64_t* __restrict__ t, int b) {
    if (b == 0) return t[0] == 0;
    if (b >= kLBounds) return false;
    return t[b - 1] + 1 == t[b];
}

// ---------------------------------------------------------------------------
template <int RW>
__global__ __launch_bounds__(kBlock) void lhist_kernel(const uint32_t* __restrict__ src, SegDev seg,
                                                       int64_t ntiles, int64_t tpb, int tile_recs, LKey lk,
                                                       const uint64_t* __restrict__ tables, int64_t tstride,
                                                       uint32_t* __restrict__ hist,
                                                       uint16_t* __restrict__ tile_cnt,
                                                       long long* __restrict__ checksum,
                                                       uint16_t* __restrict__ bkt) {
    __shared__ uint32_t s_hist[kLBins];  // the block's counts for its current segment
    __shared__ uint32_t s_th[kLBins];    // counts of the current tile
    __shared__ uint64_t s_tab[kLBounds];
    // guide: cell c of the keys kmin + [c << gsh, (c + 1) << gsh) holds
    // buckets s_guide[c] .. s_guide[c + 1] only (kmin, kmax: the table's ends)
    __shared__ uint16_t s_guide[kGuideCells + 1];
    __shared__ long long s_sum[kWaves];
    const int tid = threadIdx.x;
    for (int i = tid; i < kLBins; i += kBlock) {
        s_hist[i] = 0;
        s_th[i] = 0;
    }
    uint64_t kmin = 0, krange = 0;
    int gsh = 0;
    const int64_t t0 = (int64_t)blockIdx.x * tpb;
    const int64_t t1 = min(ntiles, t0 + tpb);
    long long csum = 0;
    int cur = -1;
    for (int64_t tile = t0; tile < t1; tile++) {
        const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
        if (s != cur) {
            __syncthreads();
            if (cur >= 0) {
                uint32_t* h = hist + (int64_t)cur * kLBins;
                for (int i = tid; i < kLBins; i += kBlock) {
                    const uint32_t c = s_hist[i];
                    if (c) atomicAdd(&h[i], c);
                    s_hist[i] = 0;
                }
            }
            if (cur < 0 || tstride != 0) {
                for (int i = tid; i < kLBounds; i += kBlock) s_tab[i] = tables[(int64_t)s * tstride + i];
                __syncthreads();
                kmin = s_tab[0];
                krange = s_tab[kLBounds - 1] - kmin;
                gsh = krange ? max(0, 64 - __clzll((long long)krange) - kGuideBits) : 0;
                for (int c = tid; c <= kGuideCells; c += kBlock) {
                    const uint64_t off = (uint64_t)c << gsh;  // exact for c < kGuideCells (gsh <= 64 - kGuideBits)
                    s_guide[c] = (uint16_t)(c == kGuideCells || off > krange ? kLBounds : bucket_of(kmin + off, s_tab));
                }
                __syncthreads();
            }
            cur = s;
        }
        const int64_t t = tile - seg.tile_start[s];
        const int64_t sbeg = seg.begin[s];
        const int64_t begin = sbeg + t * tile_recs;
        const int64_t end = min(sbeg + seg.count[s], begin + (int64_t)tile_recs);
        // U records per thread at a time, all U loads in flight before the
        // first is used
        constexpr int U = 4;
        for (int64_t i0 = begin + tid; i0 < end; i0 += U * kBlock) {
            uint32_t r[U][RW];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (i0 + (int64_t)u * kBlock < end) {
                    RecIO<RW>::load(src, i0 + (int64_t)u * kBlock, r[u]);
                } else {
#pragma unroll
                    for (int w = 0; w < RW; w++) r[u][w] = 0;
                }
            }
            // the guide narrows each search to the buckets of the key's cell
            // (lo .. lo + cnt); the U lower-bound searches then advance
            // together, their reads in flight at once.  Positions past the
            // tile search for key 0 and are ignored.
            uint64_t kq[U];
            int lo[U], cnt[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                kq[u] = lkey<RW>(r[u], lk);
                const uint64_t d = kq[u] - kmin;
                const bool inside = kq[u] >= kmin && d <= krange;
                const int c = inside ? (int)(d >> gsh) : 0;  // < kGuideCells
                const int g0 = s_guide[c], g1 = s_guide[c + 1];
                lo[u] = inside ? g0 : (kq[u] < kmin ? 0 : kLBounds);
                cnt[u] = inside ? g1 - g0 : 0;
            }
            for (;;) {
                bool more = false;
#pragma unroll
                for (int u = 0; u < U; u++) more |= cnt[u] > 0;
                if (!more) break;
                uint64_t v[U];
#pragma unroll
                for (int u = 0; u < U; u++) v[u] = s_tab[min(lo[u] + (cnt[u] >> 1), kLBounds - 1)];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int half = cnt[u] >> 1;
                    if (cnt[u] > 0) {
                        if (v[u] < kq[u]) {
                            lo[u] += half + 1;
                            cnt[u] -= half + 1;
                        } else {
                            cnt[u] = half;
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t i = i0 + (int64_t)u * kBlock;
                if (i < end) {
                    if (checksum) {
#pragma unroll
                        for (int w = 0; w < RW; w++) csum += sbyte_sum(r[u][w]);
                    }
                    const uint32_t b = (uint32_t)lo[u];  // bucket_of(key, s_tab)
                    bkt[i] = (uint16_t)b;  // for the scatter
                    atomicAdd(&s_th[b], 1u);
                }
            }
        }
        __syncthreads();
        // the tile's counts (for the scatter's offsets), added to the segment
        for (int i = tid; i < kLBins; i += kBlock) {
            const uint32_t c = s_th[i];
            tile_cnt[tile * kLBins + i] = (uint16_t)c;
            s_hist[i] += c;
            s_th[i] = 0;
        }
        __syncthreads();
    }
    if (cur >= 0) {
        uint32_t* h = hist + (int64_t)cur * kLBins;
        for (int i = tid; i < kLBins; i += kBlock) {
            const uint32_t c = s_hist[i];
            if (c) atomicAdd(&h[i], c);
        }
    }
    if (checksum) {
        csum = wave_sum_ll(csum);
        if ((tid & 63) == 0) s_sum[tid >> 6] = csum;
        __syncthreads();
        if (tid == 0) {
            long long a = 0;
            for (int w = 0; w < kWaves; w++) a += s_sum[w];
            atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
        }
    }
}

// exclusive scan of B values, one per thread of a B-thread block
template <int B>
__device__ __forceinline__ uint32_t scan_block(uint32_t v, uint32_t* s_tmp /*B / 64*/) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < B / 64; w++) pre += (w < wave) ? s_tmp[w] : 0u;
    __syncthreads();
    return pre + x - v;
}
__device__ __forceinline__ uint32_t scan1024(uint32_t v, uint32_t* s_tmp /*16*/) { return scan_block<1024>(v, s_tmp); }

// per segment: absolute base of every bucket (nb buckets per segment, a
// multiple of 1024: nb / 1024 consecutive buckets per thread)
__global__ __launch_bounds__(1024) void lscan_kernel(const uint32_t* __restrict__ hist,
                                                     const int64_t* __restrict__ seg_begin,
                                                     int64_t* __restrict__ base, int nb) {
    __shared__ uint32_t s_tmp[16];
    const int s = blockIdx.x, tid = threadIdx.x, per = nb / 1024;  // 1 .. 4 buckets per thread
    const uint32_t* h = hist + (int64_t)s * nb + (int64_t)tid * per;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        c[q] = q < per ? h[q] : 0u;
        sum += c[q];
    }
    uint32_t ex = scan1024(sum, s_tmp);
    int64_t* bs = base + (int64_t)s * nb + (int64_t)tid * per;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q < per) bs[q] = seg_begin[s] + (int64_t)ex;
        ex += c[q];
    }
}

// offset of every (tile, bucket) run inside its segment: the bucket's start
// plus the counts of the segment's earlier tiles.  Thread per (segment,
// bucket); the segment's tiles in order.
__global__ __launch_bounds__(256) void ltile_scan_kernel(const uint16_t* __restrict__ tile_cnt, SegDev seg,
                                                         const int64_t* __restrict__ bucket_base,
                                                         uint32_t* __restrict__ tile_off) {
    constexpr int PARTS = kLBins / 256;
    const int s = blockIdx.x / PARTS;
    const int b = (blockIdx.x % PARTS) * 256 + threadIdx.x;
    const int64_t t1 = seg.tile_start[s + 1];
    uint32_t run = (uint32_t)(bucket_base[(int64_t)s * kLBins + b] - seg.begin[s]);
    int64_t t = seg.tile_start[s];
    // 16 tiles' counts in flight at once per thread (one workgroup per CU:
    // the loads' latency, not bandwidth, bounds this pass)
    for (; t + 16 <= t1; t += 16) {
        uint32_t c[16];
#pragma unroll
        for (int k = 0; k < 16; k++) c[k] = tile_cnt[(t + k) * kLBins + b];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            tile_off[(t + k) * kLBins + b] = run;
            run += c[k];
        }
    }
    for (; t < t1; t++) {
        const uint32_t c = tile_cnt[t * kLBins + b];
        tile_off[t * kLBins + b] = run;
        run += c;
    }
}

// ---------------------------------------------------------------------------
// the bucket scatter: a onesweep pass with 1024 bins.  BLOCK = 512: a tile
// fits 80 KiB of LDS, so two workgroups share a CU.
template <int RW, int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK, 2048 / BLOCK) void lscatter_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, SegDev seg, int64_t ntiles,
    const uint16_t* __restrict__ bkt, const uint64_t* __restrict__ tables, int64_t tstride,
    const uint32_t* __restrict__ tile_off, uint32_t* __restrict__ error_flag, uint32_t* __restrict__ eq_dst) {
    constexpr int NW = BLOCK / 64, BPT = kLBins / BLOCK;  // waves, buckets per thread
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WAVE_TILE = 64 * ITEMS;
    static_assert(kLBins % BLOCK == 0, "buckets per thread");
    __shared__ __attribute__((aligned(16))) uint32_t s_rec[TILE * RW];
    __shared__ uint16_t s_dig[TILE];
    __shared__ uint16_t s_whist[NW][kLBins];
    __shared__ uint32_t s_excl[kLBins];
    __shared__ int32_t s_gdelta[kLBins];  // output position of the tile's run of a bucket, less its
                                          // staged position, from the segment's start
    __shared__ uint32_t s_single[kLBins / 32];
    __shared__ uint32_t s_tmp[NW];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < NW * kLBins; i += BLOCK) (&s_whist[0][0])[i] = 0;
    // workgroups go to the 8 XCDs round robin; XCD x takes the contiguous
    // tiles [x * per, (x + 1) * per), so runs of neighbouring tiles meet in
    // one L2 (the grid is 8 * per; the padding workgroups exit)
    const int64_t per = gridDim.x / 8;
    const int64_t tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
    const int64_t t = tile - seg.tile_start[s];
    const int64_t sbeg = seg.begin[s];
    const int64_t send = sbeg + seg.count[s];
    const int64_t begin = sbeg + t * TILE;
    const int64_t end = min(send, begin + (int64_t)TILE);
    const int count = (int)(end - begin);

    // records and their buckets (computed by the histogram pass)
    uint32_t r[ITEMS][RW];
    uint32_t bk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + wave * WAVE_TILE + k * 64 + lane;
        bk[k] = kLBins;
        if (i < end) {
            RecIO<RW>::load(src, i, r[k]);
            bk[k] = bkt[i];
        }
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int d = q * BLOCK + tid;
        const uint64_t m = __ballot(eq_dst != dst && single_bucket(tables + (int64_t)s * tstride, d));
        if (lane == 0) {
            s_single[d / 32] = (uint32_t)m;
            s_single[d / 32 + 1] = (uint32_t)(m >> 32);
        }
    }
    __syncthreads();  // every wave's counters are cleared by threads of all waves

    uint32_t dl[ITEMS];  // bucket | wave rank << 16
    volatile uint16_t* wh = s_whist[wave];
    const uint64_t lt_mask = __lanemask_lt();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + wave * WAVE_TILE + k * 64 + lane;
        const bool valid = i < end;
        const uint32_t d = valid ? bk[k] : (uint32_t)kLBins;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 10; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const uint32_t cnt = __popcll(peers);
        uint32_t base = 0;
        if (valid) base = wh[d];
        dl[k] = d | ((base + rank) << 16);
        if (valid && rank == 0) wh[d] = (uint16_t)(base + cnt);
    }
    __syncthreads();

    // buckets BPT * tid .. BPT * tid + BPT - 1 per thread
    uint32_t total[BPT], sum = 0;
    uint32_t toff[BPT];
#pragma unroll
    for (int q = 0; q < BPT; q++) toff[q] = tile_off[tile * kLBins + BPT * tid + q];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int d = BPT * tid + q;
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t c = s_whist[w][d];
            s_whist[w][d] = (uint16_t)acc;
            acc += c;
        }
        total[q] = acc;
        sum += acc;
    }
    uint32_t excl = scan_block<BLOCK>(sum, s_tmp);
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int d = BPT * tid + q;
        s_excl[d] = excl;
        s_gdelta[d] = (int32_t)toff[q] - (int32_t)excl;
        excl += total[q];
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t dd = dl[k] & 0xffffu;
        if (dd < kLBins) {
            const uint32_t pos = s_excl[dd] + s_whist[wave][dd] + (dl[k] >> 16);
#pragma unroll
            for (int w = 0; w < RW; w++) s_rec[pos * RW + w] = r[k][w];
            s_dig[pos] = (uint16_t)dd;
        }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int i = k * BLOCK + tid;
        if (i < count) {
            uint32_t rr[RW];
#pragma unroll
            for (int w = 0; w < RW; w++) rr[w] = s_rec[i * RW + w];
            const uint32_t dd = s_dig[i];
            const int64_t o = sbeg + (int64_t)s_gdelta[dd] + i;
            if (o < sbeg || o >= send) {
                atomicOr(error_flag, 4u);
                continue;
            }
            // a bucket of one key value is final after the stable scatter:
            // it goes straight to the sorted output
            RecIO<RW>::store(((s_single[dd >> 5] >> (dd & 31)) & 1u) ? eq_dst : dst, o, rr);
        }
    }
}

// ---------------------------------------------------------------------------
// in-LDS sorts of the buckets.  Records of a bucket are numbered wave-major
// (index = wave * 512 + item * 64 + lane) so that the per-wave counting passes
// below keep that order among equal digits: both sorts are stable.
constexpr int kCBlock = 1024, kCWaves = 16, kCItems = kChunkCap / kCBlock, kCWspan = kChunkCap / kCWaves;
constexpr int kRunMax = 512;  // largest bin ranked directly

template <int RW>
__device__ __forceinline__ void load_bucket(const uint32_t* __restrict__ src, int64_t base, int count,
                                            const LKey& lk, uint32_t (&r)[kCItems][RW],
                                            uint64_t (&kk)[kCItems], uint64_t* s_red, uint64_t& mn,
                                            uint64_t& mx) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    mn = ~0ull;
    mx = 0;
    // every load issued before any is used (indices past the bucket read its
    // last record again and are ignored)
#pragma unroll
    for (int k = 0; k < kCItems; k++) RecIO<RW>::load(src, base + min(wave * kCWspan + k * 64 + lane, count - 1), r[k]);
#pragma unroll
    for (int k = 0; k < kCItems; k++) {
        const bool valid = wave * kCWspan + k * 64 + lane < count;
        const uint64_t key = lkey<RW>(r[k], lk);
        kk[k] = valid ? key : 0;
        mn = valid && key < mn ? key : mn;
        mx = valid && key > mx ? key : mx;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t a = __shfl_xor(mn, off, 64), c = __shfl_xor(mx, off, 64);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
    }
    if (lane == 0) { s_red[wave] = mn; s_red[kCWaves + wave] = mx; }
    __syncthreads();
    mn = s_red[0];
    mx = s_red[kCWaves];
#pragma unroll
    for (int w = 1; w < kCWaves; w++) {
        mn = s_red[w] < mn ? s_red[w] : mn;
        mx = s_red[kCWaves + w] > mx ? s_red[kCWaves + w] : mx;
    }
}

template <int W = kCWaves>
__device__ __forceinline__ void block_checksum_add(long long csum, long long* s_sum, long long* checksum) {
    csum = wave_sum_ll(csum);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = csum;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long a = 0;
        for (int w = 0; w < W; w++) a += s_sum[w];
        atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
    }
}

// The receiver-table entries that lie in FirstSort bucket eb of segment ej:
// entry e gets the bucket's position in its segment (rel) plus count_le(key).
// The caller marks the bucket done.
template <int BLOCK, class CountLe>
__device__ __forceinline__ void write_entries(const EntryOut& eo, int ej, int eb, int64_t base, CountLe count_le) {
    const uint32_t* rg = eo.range + (int64_t)ej * (kLBins + 1);
    const uint32_t e0 = rg[eb], e1 = rg[eb + 1];
    const int64_t rel = base - eo.seg_begin[ej];
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += BLOCK) {
        const uint32_t r = e / kMergeBounds, k = e % kMergeBounds + 1;
        eo.bnd[merge_bnd_at(r, ej, k, eo.P)] = (int32_t)(rel + count_le(eo.keys[e]));
    }
    if (threadIdx.x == 0) eo.done[(int64_t)ej * kLBins + eb] = 1;
}

// true when the block's keys (kmn, kmx: each thread's smallest and largest,
// ~0 and 0 for none) hold one value; s_mm: 2 W words.  Contains one barrier;
// s_mm may be written again after the next barrier.
template <int W>
__device__ __forceinline__ bool block_one_value(uint64_t kmn, uint64_t kmx, uint64_t* s_mm) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t a = __shfl_xor(kmn, off, 64), b = __shfl_xor(kmx, off, 64);
        kmn = a < kmn ? a : kmn;
        kmx = b > kmx ? b : kmx;
    }
    if ((threadIdx.x & 63) == 0) {
        s_mm[threadIdx.x >> 6] = kmn;
        s_mm[W + (threadIdx.x >> 6)] = kmx;
    }
    __syncthreads();
    uint64_t mn = ~0ull, mx = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        mn = s_mm[w] < mn ? s_mm[w] : mn;
        mx = s_mm[W + w] > mx ? s_mm[W + w] : mx;
    }
    return mn == mx;
}

// One stable counting pass over 256 digits in LDS.  The records are listed
// in s_in[0..count) (identity when s_in is null); digit(e) < 256 is the digit
// of record e; the new order goes to s_out.  Positions are taken wave-major
// (wave * kCWspan + item * 64 + lane), the order in which the per-wave
// counters are laid out, so equal digits keep their order.
template <class Digit>
__device__ __forceinline__ void lds_count_pass(const uint16_t* s_in, uint16_t* s_out, int count, Digit digit,
                                               uint16_t (*s_whist)[256], uint32_t* s_excl, uint32_t* s_tmp) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < kCWaves * 256; i += kCBlock) (&s_whist[0][0])[i] = 0;
    __syncthreads();
    const uint64_t lt_mask = __lanemask_lt();
    volatile uint16_t* wh = s_whist[wave];
    uint32_t dl[kCItems];
    uint16_t el[kCItems];
#pragma unroll
    for (int k = 0; k < kCItems; k++) {
        const int pos = wave * kCWspan + k * 64 + lane;
        const bool valid = pos < count;
        const uint16_t e = valid ? (s_in ? s_in[pos] : (uint16_t)pos) : (uint16_t)0;
        const uint32_t d = valid ? digit(e) : 256u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bb = 0; bb < 8; bb++) {
            const bool bit = (d >> bb) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const uint32_t c = __popcll(peers);
        uint32_t bs = 0;
        if (valid) bs = wh[d];
        dl[k] = d | ((bs + rank) << 16);
        el[k] = e;
        if (valid && rank == 0) wh[d] = (uint16_t)(bs + c);
    }
    __syncthreads();
    uint32_t tot = 0;
    if (tid < 256) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kCWaves; w++) {
            const uint32_t c = s_whist[w][tid];
            s_whist[w][tid] = (uint16_t)acc;
            acc += c;
        }
        tot = acc;
    }
    {
        uint32_t x = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63 && wave < 4) s_tmp[wave] = x;
        __syncthreads();
        uint32_t pre = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) pre += (w < wave) ? s_tmp[w] : 0u;
        if (tid < 256) s_excl[tid] = pre + x - tot;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kCItems; k++) {
        const uint32_t d = dl[k] & 0xffffu;
        if (d < 256u) s_out[s_excl[d] + s_whist[wave][d] + (dl[k] >> 16)] = el[k];
    }
    __syncthreads();
}

// Per bucket: single-value buckets already written (eq_done) only add their
// checksum; oversized ones go to the overflow list; a bucket of one key value
// is copied.  Otherwise each record gets a bin, non-decreasing in the key:
//   plain:    d = min(FINE-1, floor(float((key - min) >> sh) * FINE / (((max - min) >> sh) + 1)))
//             (sh brings the offsets to 32 bits)
// and, when a plain bin holds more than kInterpMax records (clustered keys),
//   sampled:  64 keys sampled from the bucket and sorted split the key range
//             into 65 intervals of about equal counts, each cut into SUB bins:
//             j = #{samples <= key}, [lo, hi] = interval j,
//             d = SUB j + min(SUB-1, floor(float((key - lo) >> sh_j) * SUB / (((hi - lo) >> sh_j) + 1)))
// Records are grouped by bin with LDS atomics (in any order inside a bin),
// and a record's final slot is its bin start + #{records of the bin with a
// smaller key, or an equal key and a smaller position}: the stable order.
// In the 9-record forms, a bin of more than kLongMin records that all hold
// one key value (a duplicate-heavy key) is placed by counting instead:
// slot = bin start + #{its records before this one in the bucket}, item by
// item; the 512 x 6 form, whose registers are all in use, hands a bucket with
// such bins to the 1024-thread form (result 2) when the sample holds a
// repeated key.
// Buckets with a bin above kRunMax records of several key values go to the
// list for lsd_sort_kernel (result 1).  Record i of the bucket is item
// i / BLOCK of thread i % BLOCK.
constexpr int kSamp = 64;
constexpr int kInterpMax = 32;
constexpr int kLongMin = 32;
// Workgroup shapes of the in-LDS sort: 1024 threads x 9 records (one
// workgroup per CU) or 512 threads x 9 records (two share a CU), in 4095
// bins (65 sample intervals x 63), or 512 threads x 6 records (three share a
// CU) in 2015 bins (65 x 31).  ITEMS records per thread (CAP = BLOCK *
// ITEMS).  The bin arrays take BW words of LDS: NB u16 bin starts, then NB
// u16 bin counters, two to a word.
template <int BLOCK, int ITEMS = kCItems>
struct BSort {
    static constexpr int B = BLOCK, W = BLOCK / 64, I = ITEMS, CAP = BLOCK * ITEMS;
    static constexpr int NL = CAP / (kLongMin + 1) + 1;  // bins of more than kLongMin records, at most
    static constexpr int SUB = CAP > 3072 ? 63 : 31;             // bins per sample interval
    static constexpr int FINE = (kSamp + 1) * SUB;               // bins
    static constexpr int NB = (FINE + 1 + BLOCK - 1) / BLOCK * BLOCK;  // bin counters
    static constexpr int BW = NB;                                // words of the bin arrays
    static constexpr int TPS = BLOCK / kSamp;                    // threads per sample in the sample sort
    static constexpr int RU = ITEMS >= 8 && BLOCK <= 512 ? 8 : 4;  // keys per LDS read group in the ranking
    static_assert(NB % 2 == 0 && kSamp % TPS == 0 && CAP < 65536, "in-LDS sort shape");
};

// the records of one bucket, contiguous in src from base
template <int RW>
struct RunSrc {
    const uint32_t* recs;
    int64_t base;
    __device__ __forceinline__ void load(int i, uint32_t (&r)[RW]) const { RecIO<RW>::load(recs, base + i, r); }
};

// the records of one receiver bucket gathered from the senders' sorted
// segments, in sender order: sub-run j starts at begin[j], and records
// [pref[j], pref[j+1]) of the bucket come from it (LDS tables, nrun runs)
template <int RW>
struct GatherSrc {
    const uint32_t* recs;
    const int64_t* begin;
    const int* pref;  // 64 entries: pref[nrun .. 63] lie above every position
    int nrun;
    __device__ __forceinline__ void load(int i, uint32_t (&r)[RW]) const {
        const int j = upper_index<int>(pref, nrun, i);
        RecIO<RW>::load(recs, begin[j] + (i - pref[j]), r);
    }
};

// records idx[0..N) of src into r[0..N)
template <int RW, int N, class Src>
__device__ __forceinline__ void load_items(const Src& src, const int (&idx)[N], uint32_t (&r)[N][RW]) {
#pragma unroll
    for (int k = 0; k < N; k++) src.load(idx[k], r[k]);
}

// the gathered form: the N sub-run searches advance level by level, so each
// level's N table reads are in flight together
template <int RW, int N>
__device__ __forceinline__ void load_items(const GatherSrc<RW>& src, const int (&idx)[N], uint32_t (&r)[N][RW]) {
    int j[N];
#pragma unroll
    for (int k = 0; k < N; k++) j[k] = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        int v[N];
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = src.pref[j[k] + step];
#pragma unroll
        for (int k = 0; k < N; k++) j[k] = v[k] <= idx[k] ? j[k] + step : j[k];
    }
#pragma unroll
    for (int k = 0; k < N; k++) RecIO<RW>::load(src.recs, src.begin[j[k]] + (idx[k] - src.pref[j[k]]), r[k]);
}

// Sorts one bucket of count (1 .. BSort<BLOCK>::CAP) records, given by src,
// into dst[base .. base + count) with a BLOCK-thread workgroup.  Returns 0
// when sorted; otherwise writes nothing and returns 1 when a bin of several
// key values is crowded beyond kRunMax, 2 when the bucket is for the
// 1024-thread form (one-value long bins in the 512 x 6 form, see above).
// s_bins: BSort<BLOCK, ITEMS>::BW words of LDS owned by the caller, who may
// keep data there that src reads while loading.  When
// bounded, every key lies in [lo, hi] (the bucket's boundaries) and the
// bucket's minimum and maximum are not computed.
template <int RW, int BLOCK, class Src, int ITEMS = kCItems>
__device__ __forceinline__ int sort_bucket_lds(const Src& src, uint32_t* __restrict__ dst, int64_t base, int count,
                                                const LKey& lk, uint32_t* s_bins, long long* __restrict__ checksum_out,
                                                bool bounded = false, uint64_t lo = 0, uint64_t hi = ~0ull,
                                                const EntryOut eo = EntryOut{}, int ej = 0, int eb = 0) {
    using C = BSort<BLOCK, ITEMS>;
    // one-value long bins placed here (the 512 x 6 form hands them on)
    constexpr bool DUPS = BLOCK == kCBlock || ITEMS >= 9;
    __shared__ uint64_t s_gkey[C::CAP + 8];  // key of each bin-order slot (+8: unrolled reads)
    __shared__ uint16_t s_sbin[C::CAP];      // bin of each bin-order slot, then the final
                                             // slot of each bucket position (s_pos)
    __shared__ uint16_t s_ord[C::CAP];       // bucket position of each bin-order slot
    __shared__ uint64_t s_samp[kSamp];    // the sample, in bucket order
    __shared__ uint64_t s_sorted[kSamp];  // the sample, sorted
    __shared__ int s_crowd;
    __shared__ uint64_t s_ilo[kSamp + 1];  // sampled bins: interval j starts at key s_ilo[j],
    __shared__ uint32_t s_ish[kSamp + 1];  // its offsets are shifted right by s_ish[j]
    __shared__ float s_isc[kSamp + 1];     // and scaled by s_isc[j] into SUB bins
    __shared__ uint64_t s_red[2 * C::W];
    __shared__ uint32_t s_tmp[C::W];
    __shared__ uint32_t s_max[C::W];
    __shared__ long long s_sum[C::W];
    __shared__ int s_long;
    __shared__ int s_dup;
    __shared__ int s_defer;
    __shared__ uint32_t s_nlong;
    uint16_t* s_start = reinterpret_cast<uint16_t*>(s_bins);  // NB u16
    uint32_t* s_fill = s_bins + C::NB / 2;                    // NB u16 counters, two per word
    uint16_t* s_pos = s_sbin;
    // one-value long bins; each array reuses LDS that is free by then
    static_assert(C::NB <= 32 * 2 * kSamp && (C::W + 1) * C::NL <= C::CAP && C::CAP < 0x8000 && C::NB < 0x8000,
                  "aliased arrays of the one-value long bins");
    uint32_t* s_mix = reinterpret_cast<uint32_t*>(s_samp);   // after the sample sort: long bins of several values
    uint16_t* s_lidx = reinterpret_cast<uint16_t*>(s_fill);  // after the ranking: long bin numbers
    uint16_t* s_wc = s_ord;                 // after the ranking: their records per wave, one item
    uint16_t* s_run = s_ord + C::W * C::NL;  // and per long bin, the items before
    // one more record in bin d: the count before it
    auto bump = [&](uint32_t d) -> uint32_t {
        const uint32_t sh = (d & 1u) << 4;
        return (atomicAdd(&s_fill[d >> 1], 1u << sh) >> sh) & 0xffffu;
    };

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    // Every load is issued before any of them is used, so they are all in
    // flight at once: item k of thread t is record k * BLOCK + t (indices
    // past the bucket read its last record again and are ignored); thread
    // q < 64 also reads the sample key of record floor(q * count / 64).
    // r[C::I]: the sample record
    uint32_t r[C::I + 1][RW];
    {
        int idx[C::I + 1];
#pragma unroll
        for (int k = 0; k < C::I; k++) idx[k] = min(k * C::B + tid, count - 1);
        idx[C::I] = ((tid & (kSamp - 1)) * count) >> 6;
        load_items<RW>(src, idx, r);
    }
    uint64_t kk[C::I];
    uint64_t mn = ~0ull, mx = 0;
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        const bool valid = k * C::B + tid < count;
        const uint64_t key = lkey<RW, (ITEMS >= 9)>(r[k], lk);
        kk[k] = valid ? key : 0;
        mn = valid && key < mn ? key : mn;
        mx = valid && key > mx ? key : mx;
    }
    if (tid < kSamp) s_samp[tid] = lkey<RW, (ITEMS >= 9)>(r[C::I], lk);
    if (tid == 0) {
        s_crowd = 0;
        s_long = 0;
        s_dup = 0;
        s_defer = 0;
        s_nlong = 0;
    }
    if (bounded) {
        mn = lo;
        mx = hi;
        __syncthreads();  // every load of src is done: s_bins is free
    } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t a = __shfl_xor(mn, off, 64), c = __shfl_xor(mx, off, 64);
            mn = a < mn ? a : mn;
            mx = c > mx ? c : mx;
        }
        if (lane == 0) { s_red[wave] = mn; s_red[C::W + wave] = mx; }
        __syncthreads();  // every load of src is done: s_bins is free
        mn = s_red[0];
        mx = s_red[C::W];
#pragma unroll
        for (int w = 1; w < C::W; w++) {
            mn = s_red[w] < mn ? s_red[w] : mn;
            mx = s_red[C::W + w] > mx ? s_red[C::W + w] : mx;
        }
    }
    long long csum = 0;
    if (mn == mx) {
        // one key value: the stable order is the current order
#pragma unroll
        for (int k = 0; k < C::I; k++) {
            const int i = k * C::B + tid;
            if (i < count) {
                RecIO<RW>::store(dst, base + i, r[k]);
                if (checksum_out) {
#pragma unroll
                    for (int w = 0; w < RW; w++) csum += sbyte_sum(r[k][w]);
                }
            }
        }
        if (checksum_out) block_checksum_add<C::W>(csum, s_sum, checksum_out);
        if (eo.keys) {
            const uint64_t v = mn;
            write_entries<BLOCK>(eo, ej, eb, base, [&](uint64_t t) { return t >= v ? (uint32_t)count : 0u; });
        }
        return 0;
    }
    for (int i = tid; i < C::NB / 2; i += C::B) s_fill[i] = 0;
    // plain bins: offsets from the minimum shifted to 32 bits, then scaled
    // (shift, conversion and scaling are all monotone)
    const int pbits = 64 - __clzll((long long)((mx - mn) | 1));
    const int psh = pbits > 32 ? pbits - 32 : 0;
    const float scale = (float)C::FINE / ((float)(uint32_t)((mx - mn) >> psh) + 1.0f);
    auto plain_bin = [&](uint64_t key) -> uint32_t {
        const uint32_t d = (uint32_t)((float)(uint32_t)((key - mn) >> psh) * scale);
        return d > (uint32_t)(C::FINE - 1) ? (uint32_t)(C::FINE - 1) : d;
    };
    {
        // sort the sample: TPS threads rank one sample, each against
        // kSamp / TPS others (ties by index), sums over the TPS lanes
        constexpr int PER = kSamp / C::TPS;
        const int a = tid / C::TPS, part = tid % C::TPS;
        const uint64_t v = s_samp[a];
        uint32_t c = 0;
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int q = part * PER + u;
            const uint64_t o = s_samp[q];
            c += (o < v || (o == v && q < a)) ? 1u : 0u;
        }
#pragma unroll
        for (int off = 1; off < C::TPS; off <<= 1) c += __shfl_xor(c, off, C::TPS);
        if (part == 0) s_sorted[c] = v;
    }
    __syncthreads();
    if (tid <= kSamp) {
        // interval j (keys k with s_sorted[j-1] <= k < s_sorted[j]) spans the
        // offsets 0 .. span from its first key; offsets shifted to 32 bits
        const int j = tid;
        const uint64_t lo = j == 0 ? mn : s_sorted[j - 1];
        const uint64_t hi = j == kSamp ? mx : (s_sorted[j] > lo ? s_sorted[j] - 1 : lo);
        const uint64_t span = hi - lo;
        const int bits = 64 - __clzll((long long)(span | 1));
        const int sh = bits > 32 ? bits - 32 : 0;
        s_ilo[j] = lo;
        s_ish[j] = (uint32_t)sh;
        s_isc[j] = (float)C::SUB / ((float)(uint32_t)(span >> sh) + 1.0f);
        // two neighbouring samples in one plain bin put about count / 64
        // records there: the sampled bins are used
        if (j + 1 < kSamp && plain_bin(s_sorted[j]) == plain_bin(s_sorted[j + 1])) s_crowd = 1;
        // a key sampled twice: duplicate-heavy keys are likely
        if (j + 1 < kSamp && s_sorted[j] == s_sorted[j + 1]) s_dup = 1;
    }
    for (int i = tid; i < (C::NB + 31) / 32; i += C::B) s_mix[i] = 0;  // the sample is sorted
    __syncthreads();
    int mode = count >= kInterpMax * kSamp && s_crowd ? 1 : 0;  // 0: plain bins, 1: sampled bins
    // sampled: j = #{samples <= key}, then the interval's offset shifted and
    // scaled into SUB bins
    auto bin_of = [&](uint64_t key) -> uint32_t {
        if (mode == 0) return plain_bin(key);
        int j = 0;
#pragma unroll
        for (int step = kSamp / 2; step > 0; step >>= 1)
            if (s_sorted[j + step - 1] <= key) j += step;
        if (j == kSamp - 1 && s_sorted[kSamp - 1] <= key) j = kSamp;
        const uint32_t x = (uint32_t)((key - s_ilo[j]) >> s_ish[j]);
        uint32_t sub = (uint32_t)((float)x * s_isc[j]);
        sub = sub > (uint32_t)(C::SUB - 1) ? (uint32_t)(C::SUB - 1) : sub;
        return (uint32_t)j * C::SUB + sub;
    };
    uint32_t dbin[C::I];
    uint32_t bmax = 0;  // the fullest bin
    for (;;) {
#pragma unroll
        for (int k = 0; k < C::I; k++) {
            const int i = k * C::B + tid;
            dbin[k] = 0;
            if (i < count) {
                dbin[k] = bin_of(kk[k]);
                bump(dbin[k]);
            }
        }
        __syncthreads();
        // exclusive scan of the NB counts (those past FINE zero), NB / BLOCK
        // per thread; the counters are reset for the grouping below
        constexpr int PER = C::NB / C::B;
        static_assert(PER % 2 == 0, "bin counters in pairs");
        uint32_t c[PER], sum = 0, big = 0;
#pragma unroll
        for (int q = 0; q < PER / 2; q++) {
            const uint32_t w = s_fill[tid * (PER / 2) + q];
            s_fill[tid * (PER / 2) + q] = 0;
            c[2 * q] = w & 0xffffu;
            c[2 * q + 1] = w >> 16;
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            sum += c[j];
            big = c[j] > big ? c[j] : big;
        }
        uint32_t pre = scan_block<C::B>(sum, s_tmp);
#pragma unroll
        for (int j = 0; j < PER; j++) { s_start[tid * PER + j] = pre; pre += c[j]; }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_xor(big, off, 64);
            big = o > big ? o : big;
        }
        if (lane == 0) s_max[wave] = big;
        __syncthreads();
        bmax = 0;
#pragma unroll
        for (int w = 0; w < C::W; w++) bmax = s_max[w] > bmax ? s_max[w] : bmax;
        // clustered keys the sample did not reveal: sampled bins instead
        if (mode == 1 || bmax <= (uint32_t)kInterpMax) break;
        mode = 1;
    }
    // one-value long bins are looked for when the sample showed a repeated key
    const bool any_long = s_dup && bmax > (uint32_t)kLongMin;
    if (!DUPS && any_long) return 2;
    // group by bin: key, bin and bucket position of every bin-order slot
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        const int i = k * C::B + tid;
        if (i < count) {
            const uint32_t o = s_start[dbin[k]] + bump(dbin[k]);
            s_gkey[o] = kk[k];
            s_sbin[o] = (uint16_t)dbin[k];
            s_ord[o] = (uint16_t)i;
        }
    }
    __syncthreads();
    // final slots, one bin-order slot per thread (threads of a wave mostly
    // share a bin, so the loops below stay coherent).  The keys of a bin are
    // read RU at a time; the up to RU - 1 read past its end are keys of
    // later bins (larger) or the padding (all ones), which add no smaller
    // key.  Equal keys keep their bucket order.
    if (tid < 8) s_gkey[count + tid] = ~0ull;
    if (DUPS && any_long) {
        // a long bin with a slot whose key is not its first slot's holds
        // several key values
#pragma unroll
        for (int k = 0; k < C::I; k++) {
            const int o = k * C::B + tid;
            if (o < count) {
                const uint32_t d = s_sbin[o], a = s_start[d];
                if (s_start[d + 1] - a > (uint32_t)kLongMin && s_gkey[o] != s_gkey[a])
                    atomicOr(&s_mix[d >> 5], 1u << (d & 31));
            }
        }
    }
    __syncthreads();
    // the slots' bins, positions, keys and bin bounds are read for all items
    // before the first is ranked, so the reads are in flight together (bin
    // and position, start and end, are below 2^16: two to a register)
    uint32_t rb[C::I], ri[C::I], rs0[C::I], re[C::I];
    uint64_t rk[C::I];
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        const int o = min(k * C::B + tid, count - 1);
        rb[k] = s_sbin[o];
        ri[k] = s_ord[o];
        rk[k] = s_gkey[o];
    }
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        rs0[k] = s_start[rb[k]];
        re[k] = s_start[rb[k] + 1];
    }
    __syncthreads();  // every slot's bin is read: s_sbin becomes s_pos
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        if (k * C::B + tid >= count) continue;
        const uint32_t i = ri[k], b = rb[k];
        const uint32_t s0 = rs0[k], e = re[k];
        if (e - s0 == 1) {
            s_pos[i] = (uint16_t)s0;
        } else if (DUPS && any_long && e - s0 > (uint32_t)kLongMin && !((s_mix[b >> 5] >> (b & 31)) & 1u)) {
            // one key value: placed by counting below; s_pos holds 0x8000 | bin
            s_pos[i] = (uint16_t)(0x8000u | b);
            s_defer = 1;
        } else if (e - s0 > (uint32_t)kRunMax) {
            s_long = 1;
        } else {
            const uint64_t key = rk[k];
            uint32_t lt = 0, eq = 0;
            for (uint32_t j = s0; j < e; j += C::RU) {
                uint64_t k4[C::RU];
#pragma unroll
                for (int u = 0; u < C::RU; u++) k4[u] = s_gkey[j + u];
#pragma unroll
                for (int u = 0; u < C::RU; u++) {
                    lt += k4[u] < key ? 1u : 0u;
                    eq += k4[u] == key ? 1u : 0u;
                }
            }
            if (eq > 1) {
                // equal keys (or the padding): the exact count of the equal
                // keys earlier in the bucket
                eq = 0;
                for (uint32_t j = s0; j < e; j++) eq += (s_gkey[j] == key && s_ord[j] < i) ? 1u : 0u;
                lt += eq;
            }
            s_pos[i] = (uint16_t)(s0 + lt);
        }
    }
    __syncthreads();
    if (s_long) return 1;
    if (DUPS && s_defer) {
        // the records of the one-value long bins, item by item (records
        // k B .. k B + B - 1, wave by wave): slot = bin start + the bin's
        // records in earlier items + in earlier waves of this item + in
        // earlier lanes of this wave.  Long bins are numbered first (in any
        // order).
        {
            constexpr int PER = C::NB / C::B;
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int d = tid * PER + j;
                const bool lg = d < C::FINE && s_start[d + 1] - s_start[d] > (uint32_t)kLongMin;
                s_lidx[d] = lg ? (uint16_t)atomicAdd(&s_nlong, 1u) : (uint16_t)0xffffu;
            }
            for (int q = tid; q < C::W * C::NL; q += C::B) s_wc[q] = 0;
            for (int q = tid; q < C::NL; q += C::B) s_run[q] = 0;
        }
        __syncthreads();
        for (int k = 0; k < C::I; k++) {
            const int i = k * C::B + tid;
            // this wave's records of each long bin in item k (the lanes of
            // a bin found by ballots, one bin at a time)
            {
                const uint32_t sp = i < count ? s_pos[i] : 0u;
                const bool def = (sp & 0x8000u) != 0;
                const uint32_t L = def ? s_lidx[sp & 0x7fffu] : 0xffffu;
                uint64_t act = __ballot(def);
                while (act) {
                    const int ld = __builtin_ctzll(act);
                    const uint32_t Ll = (uint32_t)__builtin_amdgcn_readlane((int)L, ld);
                    const uint64_t m = __ballot(L == Ll);
                    if (lane == ld) s_wc[wave * C::NL + Ll] = (uint16_t)__popcll(m);
                    act &= ~m;
                }
            }
            __syncthreads();
            {
                const uint32_t sp = i < count ? s_pos[i] : 0u;
                const bool def = (sp & 0x8000u) != 0;
                const uint32_t d = sp & 0x7fffu;
                const uint32_t L = def ? s_lidx[d] : 0xffffu;
                uint32_t before = 0;
                if (def) {
                    before = s_start[d] + s_run[L];
                    for (int w = 0; w < wave; w++) before += s_wc[w * C::NL + L];
                }
                uint64_t act = __ballot(def);
                while (act) {
                    const int ld = __builtin_ctzll(act);
                    const uint32_t Ll = (uint32_t)__builtin_amdgcn_readlane((int)L, ld);
                    const uint64_t m = __ballot(L == Ll);
                    if (L == Ll) s_pos[i] = (uint16_t)(before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)));
                    act &= ~m;
                }
            }
            __syncthreads();
            for (int q = tid; q < C::NL; q += C::B) {
                uint32_t acc = 0;
#pragma unroll
                for (int w = 0; w < C::W; w++) {
                    acc += s_wc[w * C::NL + q];
                    s_wc[w * C::NL + q] = 0;
                }
                s_run[q] = (uint16_t)(s_run[q] + acc);
            }
            __syncthreads();
        }
    }
    if (eo.keys) {
        // #{keys <= t}: every key of the bins before t's bin, and those of
        // its bin that are <= t (bins are non-decreasing in the key; the keys
        // lie in [mn, mx])
        write_entries<BLOCK>(eo, ej, eb, base, [&](uint64_t t) -> uint32_t {
            if (t < mn) return 0u;
            if (t >= mx) return (uint32_t)count;
            const uint32_t d = bin_of(t);
            const uint32_t a = s_start[d], e = s_start[d + 1];
            // a long bin of one key value (duplicate-heavy keys, where many
            // entries fall): all of it or none, without reading it
            if (DUPS && any_long && e - a > (uint32_t)kLongMin && !((s_mix[d >> 5] >> (d & 31)) & 1u))
                return t >= s_gkey[a] ? e : a;
            uint32_t c = a;
            for (uint32_t q = a; q < e; q++) c += s_gkey[q] <= t ? 1u : 0u;
            return c;
        });
    }
#pragma unroll
    for (int k = 0; k < C::I; k++) {
        const int i = k * C::B + tid;
        if (i < count) {
            RecIO<RW>::store(dst, base + s_pos[i], r[k]);
            if (checksum_out) {
#pragma unroll
                for (int w = 0; w < RW; w++) csum += sbyte_sum(r[k][w]);
            }
        }
    }
    if (checksum_out) block_checksum_add<C::W>(csum, s_sum, checksum_out);
    return 0;
}

// One workgroup per bucket; the bucket's count, base and table bounds are
// read together.
// BLOCK = 512: buckets above BSort<512>::CAP (up to kChunkCap) go to the
// medium list for the 1024-thread form (medium must then be given).
template <int RW, int BLOCK, int ITEMS = kCItems>
__global__ __launch_bounds__(BLOCK, ITEMS <= 6 ? 6 : 2048 / BLOCK) void interp_sort_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, LKey lk,
    const uint32_t* __restrict__ hist, const int64_t* __restrict__ bucket_base,
    const uint64_t* __restrict__ tables, int64_t tstride, int64_t nbuckets, int64_t* __restrict__ overflow,
    uint32_t* __restrict__ noverflow, int max_overflow, uint32_t* __restrict__ hard, uint32_t* __restrict__ nhard,
    uint32_t* __restrict__ medium, uint32_t* __restrict__ nmedium, long long* __restrict__ checksum_out,
    int eq_done, EntryOut eo, uint32_t ovf_cap) {
    using C = BSort<BLOCK, ITEMS>;
    __shared__ long long s_sum[C::W];
    __shared__ uint32_t s_bins[C::BW];
    const int tid = threadIdx.x;
    // count, base and the table entries t[k-1], t[k] around bucket k
    struct Meta {
        uint32_t cnt;
        int64_t base;
        uint64_t tlo, thi;
    };
    auto meta = [&](int64_t bb) -> Meta {
        Meta m{hist[bb], bucket_base[bb], 0, ~0ull};
        if (tables) {
            const int kb = (int)(bb % kLBins);
            const uint64_t* t = tables + (bb / kLBins) * tstride;
            if (kb > 0) m.tlo = t[kb - 1];
            if (kb < kLBounds) m.thi = t[kb];
        }
        return m;
    };
    {
        const int64_t b = blockIdx.x;
        if (b >= nbuckets) return;
        const Meta cur = meta(b);
        const uint32_t cnt32 = cur.cnt;
        const int64_t base = cur.base;
        const int kb = (int)(b % kLBins);
        // bucket k holds one key value (see single_bucket)
        const bool single = tables && (kb == 0 ? cur.thi == 0 : (kb < kLBounds && cur.tlo + 1 == cur.thi));
        const int ej = (int)(b / kLBins);
        const uint32_t cap = ovf_cap < (uint32_t)kChunkCap ? ovf_cap : (uint32_t)kChunkCap;
        if (cnt32 == 0) {
            if (eo.keys) write_entries<BLOCK>(eo, ej, kb, base, [](uint64_t) { return 0u; });
        } else if (eq_done && single) {
            if (eo.keys) {
                const uint64_t v = cur.thi;
                write_entries<BLOCK>(eo, ej, kb, base, [&](uint64_t t) { return t >= v ? cnt32 : 0u; });
            }
            if (checksum_out) {
                long long csum = 0;
                for (int64_t i = base + tid; i < base + (int64_t)cnt32; i += BLOCK) {
                    uint32_t rr[RW];
                    RecIO<RW>::load(dst, i, rr);
#pragma unroll
                    for (int w = 0; w < RW; w++) csum += sbyte_sum(rr[w]);
                }
                block_checksum_add<C::W>(csum, s_sum, checksum_out);
            }
        } else if (cnt32 > (uint32_t)C::CAP && cnt32 <= cap) {
            // the 1024-thread form, or without a medium list the LDS radix sort
            if (tid == 0) {
                if (medium) medium[atomicAdd(nmedium, 1u)] = (uint32_t)b;
                else hard[atomicAdd(nhard, 1u)] = (uint32_t)b;
            }
        } else if (cnt32 > cap) {
            if (tid == 0) {
                const uint32_t slot = atomicAdd(noverflow, 1u);
                if ((int)slot < max_overflow) {
                    overflow[2 * slot] = base;
                    overflow[2 * slot + 1] = (int64_t)cnt32;
                }
            }
        } else {
            __builtin_assume(cnt32 <= (uint32_t)C::CAP && cnt32 > 0u);
            const RunSrc<RW> rs{src, base};
            // an interior bucket's keys lie in (t[k-1], t[k]]
            const bool bounded = tables && kb > 0 && kb < kLBins - 1 && cur.tlo < cur.thi;
            const int rc = sort_bucket_lds<RW, BLOCK, RunSrc<RW>, ITEMS>(
                rs, dst, base, (int)cnt32, lk, s_bins, checksum_out, bounded, bounded ? cur.tlo + 1 : 0,
                bounded ? cur.thi : ~0ull, eo, ej, kb);
            if (rc == 1 && tid == 0) hard[atomicAdd(nhard, 1u)] = (uint32_t)b;
            if (rc == 2 && tid == 0) {
                if (medium) medium[atomicAdd(nmedium, 1u)] = (uint32_t)b;  // the 1024-thread form
                else hard[atomicAdd(nhard, 1u)] = (uint32_t)b;             // no medium list
            }
        }
    }
}

// The listed buckets (ids into hist / bucket_base, nbins buckets and nbounds
// table boundaries per segment), sorted in place by 1024-thread workgroups:
// the merged SecondSort's buckets above its own in-LDS capacity, already
// copied in sender order.  A crowded bin sends the bucket on to the LDS radix
// sort's list.
template <int RW>
__global__ __launch_bounds__(kCBlock) void listed_sort_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, LKey lk, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ hist, const int64_t* __restrict__ bucket_base, const uint64_t* __restrict__ tables,
    int nbins, int64_t tstride, uint32_t* __restrict__ hard, uint32_t* __restrict__ nhard,
    long long* __restrict__ checksum_out, EntryOut eo) {
    __shared__ uint32_t s_bins[BSort<kCBlock>::BW];
    const uint32_t b = list[blockIdx.x];
    const uint32_t cnt32 = hist[b];
    const int64_t base = bucket_base[b];
    const int k = (int)(b % (uint32_t)nbins);
    const uint64_t* t = tables + (int64_t)(b / (uint32_t)nbins) * tstride;
    const uint64_t tlo = k > 0 ? t[k - 1] : 0, thi = k < nbins - 1 ? t[k] : ~0ull;
    if (cnt32 == 0 || cnt32 > (uint32_t)kChunkCap) return;
    // every record is in registers before the first store: src == dst is safe
    const RunSrc<RW> rs{src, base};
    const bool bounded = k > 0 && k < nbins - 1 && tlo < thi;
    if (sort_bucket_lds<RW, kCBlock>(rs, dst, base, (int)cnt32, lk, s_bins, checksum_out, bounded,
                                     bounded ? tlo + 1 : 0, bounded ? thi : ~0ull, eo, (int)(b / (uint32_t)nbins), k))
        if (threadIdx.x == 0) hard[atomicAdd(nhard, 1u)] = b;
}

// ---------------------------------------------------------------------------
// SecondSort without the exchange.  Receiver r's records from sender j are
// positions [myC[j][r], myC[j][r+1]) of sender j's sorted segment; bucket k of
// r holds the keys in (t_r[k-1], t_r[k]], i.e. one sub-run per sender.  Every
// receiver has kMergeBins buckets (tables of kMergeBounds boundaries), three
// times as many as a FirstSort segment, because the model's boundaries at
// this resolution leave bucket sizes spread widely.  Runs of neighbouring
// buckets of a receiver are grouped greedily (up to kGroupMax buckets,
// together at most kMergeCap records: their sub-runs are adjacent in every
// sender) and sorted in one pass by a kMergeBlock-thread workgroup of
// kMergeItems records per thread (three share a CU); a bucket above
// kMergeCap is sorted alone by kSmallItems records per thread (two per CU).
//   merge_bounds_kernel  bnd[r][j][k] = start of bucket k's sub-run in sender
//                        j (relative to j's segment), k = 0..kMergeBins (one
//                        sender's bounds of one receiver side by side)
//   merge_count_kernel   bucket sizes (sum over senders)
//   merge_group_kernel   the groups, one workgroup per receiver: greedy over
//                        chunks of kGroupChunk buckets, one thread per chunk
//   merge_bucket_kernel  one workgroup per group: gathers the sub-runs in
//                        sender order (the order the exchange would have
//                        left them in) and sorts them in LDS; single-value
//                        buckets are copied; a group with a crowded bin and
//                        buckets above the LDS capacity are copied bucket by
//                        bucket in that order and listed for the LDS radix
//                        sort, the 1024-thread in-LDS sort or the radix path
constexpr int kMergeMaxP = 64;
constexpr int kGroupMax = 8;     // buckets per group (task code: bucket << 3 | count - 1)
constexpr int kGroupChunk = 48;  // groups never cross a chunk of this many buckets
constexpr int kMergeBlock = 512;
constexpr int kSmallItems = 9;  // FirstSort buckets up to 4608 records in 512-thread workgroups
constexpr int kMergeItems = 6;  // groups up to 3072 records in 512-thread workgroups, three per CU
static_assert(BSort<kMergeBlock, kMergeItems>::CAP == kMergeCap, "merged SecondSort capacity");
static_assert(BSort<kMergeBlock, kSmallItems>::CAP == kMergeBigCap, "merged SecondSort capacity");

template <int RW>
// Positions in the whole sorted segment of every sender (the receivers'
// windows are applied by merge_count_kernel, so this runs before the
// splitters are known): bnd[r][j][k] = #{keys of sender j <= t_r[k-1]} for
// k = 1 .. kMergeBins - 1, 0 for k = 0, the segment's size for kMergeBins.
// win (the windows myC, when known): each search covers receiver r's
// window in sender j only (the same value once clamped to it); with the
// receivers' even samples (samp, spref: merge_sample_tables_kernel) only
// the records between sender j's two samples that bracket the entry.
__global__ __launch_bounds__(256) void merge_bounds_kernel(const uint32_t* __restrict__ recs, LKey lk,
                                                           const int64_t* __restrict__ seg_begin,
                                                           const int64_t* __restrict__ seg_count, int P,
                                                           const uint64_t* __restrict__ tables,
                                                           const uint64_t* __restrict__ ftab, int64_t ftstride,
                                                           const int64_t* __restrict__ fbase,
                                                           const int64_t* __restrict__ win,
                                                           const uint64_t* __restrict__ samp,
                                                           const int64_t* __restrict__ spref,
                                                           int32_t* __restrict__ bnd) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (r, j, k), k = 0..kMergeBins
    if (g >= (int64_t)P * P * (kMergeBins + 1)) return;
    const int k = (int)(g % (kMergeBins + 1));
    const int64_t rj = g / (kMergeBins + 1);
    const int j = (int)(rj % P), r = (int)(rj / P);
    const int64_t cnt = seg_count[j];
    int64_t res;
    if (k == 0) {
        res = 0;
    } else if (k == kMergeBins) {
        res = cnt;
    } else {
        // first position with key > t_r[k-1]
        const uint64_t t = tables[(int64_t)r * kMergeBounds + (k - 1)];
        const int64_t sb = seg_begin[j];
        int64_t a = 0, e = cnt;
        if (ftab) {
            // it lies in FirstSort bucket m = #{sender j's table entries < t},
            // the keys in (ft[m-1], ft[m]]: every key of earlier buckets is
            // <= t, every key of later ones > t
            const uint64_t* ft = ftab + (int64_t)j * ftstride;
            int m = 0;
#pragma unroll
            for (int step = 512; step > 0; step >>= 1)
                if (m + step <= kLBounds && ft[m + step - 1] < t) m += step;
            const int64_t* fb = fbase + (int64_t)j * kLBins;
            a = fb[m] - sb;
            e = m + 1 < kLBins ? fb[m + 1] - sb : cnt;
        } else if (win) {
            const int64_t* wj = win + (int64_t)j * (P + 1);
            a = wj[r];
            e = wj[r + 1];
            if (samp && a < e) {
                // sample u lies at position pos(u) of the receiver's records;
                // sender j's samples are u0 .. u1 - 1, at window offsets
                // pos(u) - pj, keys non-decreasing
                const int64_t* pr = spref + (int64_t)r * (kMergeMaxP + 1);
                const int64_t nr = pr[P], pj = pr[j], pj1 = pr[j + 1];
                auto pos = [&](int64_t u) -> int64_t { return ((2 * u + 1) * nr) / (2 * kMergeSamples); };
                auto first_at = [&](int64_t p) -> int64_t {  // least u with pos(u) >= p
                    int64_t x = 0, y = kMergeSamples;
                    while (x < y) {
                        const int64_t m = (x + y) >> 1;
                        if (pos(m) < p) x = m + 1; else y = m;
                    }
                    return x;
                };
                const int64_t u0 = first_at(pj), u1 = first_at(pj1);
                const uint64_t* sr = samp + (int64_t)r * kMergeSamples;
                int64_t qa = u0, qe = u1;  // first of sender j's samples above t
                while (qa < qe) {
                    const int64_t m = (qa + qe) >> 1;
                    if (sr[m] <= t) qa = m + 1; else qe = m;
                }
                const int64_t w0 = a;
                if (qa > u0) a = w0 + (pos(qa - 1) - pj) + 1;
                if (qa < u1) e = w0 + (pos(qa) - pj);
            }
        }
        while (a < e) {
            const int64_t m = a + ((e - a) >> 1);
            uint32_t rr[RW];
            RecIO<RW>::load(recs, sb + m, rr);
            if (lkey<RW>(rr, lk) <= t) a = m + 1; else e = m;
        }
        res = a;
    }
    bnd[g] = (int32_t)res;  // merge_bnd_at(r, j, k, P)
}

// Receiver tables of the merged SecondSort taken from the records (for
// callers without a model): kMergeSamples keys of receiver r at evenly spaced
// positions of its records (the windows [myC[j][r], myC[j][r+1]) of the
// senders' sorted segments, in sender order), sorted in LDS (bitonic over
// kTabSort slots, the rest padded with the largest key); entry e is sample
// 4e + 3, so that a bucket holds about four samples' share of the records.
// Where entries repeat a key v, the first of them becomes v - 1: the bucket
// (v - 1, v] then holds that key alone and is copied, not sorted.  One
// 1024-thread workgroup per receiver.
constexpr int kTabSort = 16384;
static_assert(kMergeSamples <= kTabSort && (kTabSort & (kTabSort - 1)) == 0, "bitonic sort of the samples");
static_assert(4 * (kMergeBounds - 1) + 3 < kMergeSamples, "entries are samples");
// The unsorted sample (samp[r][u]: the key at position pos(u) of receiver
// r's records) and the prefixes over the senders' windows (spref[r][j], j =
// 0 .. P) are kept for merge_bounds_kernel.
template <int RW>
__global__ __launch_bounds__(1024) void merge_sample_tables_kernel(const uint32_t* __restrict__ recs, LKey lk,
                                                                   const int64_t* __restrict__ seg_begin, int P,
                                                                   const int64_t* __restrict__ myC,
                                                                   uint64_t* __restrict__ tables,
                                                                   uint64_t* __restrict__ samp,
                                                                   int64_t* __restrict__ spref) {
    __shared__ uint64_t s_k[kTabSort];
    __shared__ int64_t s_pref[kMergeMaxP + 1];  // receiver r's records in the windows before sender j's
    __shared__ int64_t s_lo[kMergeMaxP];        // first record of sender j's window (absolute)
    const int r = blockIdx.x, tid = threadIdx.x;
    if (tid < 64) {
        int64_t len = 0, lo = 0;
        if (tid < P) {
            const int64_t* cj = myC + (int64_t)tid * (P + 1);
            lo = seg_begin[tid] + cj[r];
            len = cj[r + 1] - cj[r];
        }
        int64_t x = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int64_t y = __shfl_up(x, off, 64);
            if (tid >= off) x += y;
        }
        if (tid < P) {
            s_pref[tid] = x - len;
            s_lo[tid] = lo;
            spref[(int64_t)r * (kMergeMaxP + 1) + tid] = x - len;
        }
        if (tid == P - 1) {
            s_pref[P] = x;
            spref[(int64_t)r * (kMergeMaxP + 1) + P] = x;
        }
    }
    __syncthreads();
    const int64_t nr = s_pref[P];
    for (int u = tid; u < kTabSort; u += 1024) {
        uint64_t key = ~0ull;
        if (u < kMergeSamples && nr > 0) {
            const int64_t pos = ((2 * (int64_t)u + 1) * nr) / (2 * kMergeSamples);
            // the sender whose window holds pos: the last j with s_pref[j] <= pos
            int j = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
                if (j + step < P && s_pref[j + step] <= pos) j += step;
            uint32_t rr[RW];
            RecIO<RW>::load(recs, s_lo[j] + (pos - s_pref[j]), rr);
            key = lkey<RW>(rr, lk);
        }
        s_k[u] = key;
        if (u < kMergeSamples) samp[(int64_t)r * kMergeSamples + u] = key;
    }
    for (int size = 2; size <= kTabSort; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = tid; t < kTabSort / 2; t += 1024) {
                const int i = 2 * t - (t & (stride - 1));  // lower slot of the pair
                const uint64_t a = s_k[i], b = s_k[i + stride];
                if ((a > b) == ((i & size) == 0)) {
                    s_k[i] = b;
                    s_k[i + stride] = a;
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < kMergeBounds; e += 1024) {
        const uint64_t v = s_k[4 * e + 3];
        const bool first = e == 0 || s_k[4 * e - 1] < v;
        const bool repeated = e + 1 < kMergeBounds && s_k[4 * e + 7] == v;
        tables[(int64_t)r * kMergeBounds + e] = first && repeated && v > 0 ? v - 1 : v;
    }
}

// FirstSort tables without a model: S keys of every segment at evenly
// spaced positions (out[j * S + u] = the key of record floor((2u + 1) cnt_j
// / 2S) of segment j; all ones for an empty segment), sorted by the caller;
// then entry e of segment j's table is sorted sample (e + 1) S / kLBins - 1,
// and the first of a run of equal entries v becomes v - 1, as in the
// receiver tables (bucket (v - 1, v] holds that key alone).
template <int RW>
__global__ __launch_bounds__(256) void seg_sample_kernel(const uint32_t* __restrict__ recs, LKey lk,
                                                         const int64_t* __restrict__ seg_begin,
                                                         const int64_t* __restrict__ seg_count, int P, int S,
                                                         uint64_t* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)P * S) return;
    const int j = (int)(g / S);
    const int64_t u = g % S, cnt = seg_count[j];
    uint64_t key = ~0ull;
    if (cnt > 0) {
        uint32_t rr[RW];
        RecIO<RW>::load(recs, seg_begin[j] + ((2 * u + 1) * cnt) / (2 * (int64_t)S), rr);
        key = lkey<RW>(rr, lk);
    }
    out[g] = key;
}

__global__ __launch_bounds__(256) void sample_tables_kernel(const uint64_t* __restrict__ sorted, int P, int S,
                                                            uint64_t* __restrict__ tables) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (j, e)
    if (g >= (int64_t)P * kLBounds) return;
    const int j = (int)(g / kLBounds), e = (int)(g % kLBounds);
    const int step = S / kLBins;
    const uint64_t* sj = sorted + (int64_t)j * S;
    const uint64_t v = sj[(e + 1) * step - 1];
    const bool first = e == 0 || sj[e * step - 1] < v;
    const bool repeated = e + 1 < kLBounds && sj[(e + 2) * step - 1] == v;
    tables[g] = first && repeated && v > 0 ? v - 1 : v;
}

// Receiver tables made before the splitters are known (from the model)
// keep kEdgeReserve entries at each end free: they repeat the first and last
// of the others (empty buckets) until merge_edges_kernel gives them keys.
__global__ __launch_bounds__(256) void tables_reserve_edges_kernel(uint64_t* __restrict__ tables, int P) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;  // (r, i), i < 2 kEdgeReserve
    if (g >= P * 2 * kEdgeReserve) return;
    const int r = g / (2 * kEdgeReserve), i = g % (2 * kEdgeReserve);
    uint64_t* t = tables + (int64_t)r * kMergeBounds;
    if (i < kEdgeReserve) t[i] = t[kEdgeReserve];
    else t[kMergeBounds - 2 * kEdgeReserve + i] = t[kMergeBounds - kEdgeReserve - 1];
}

// Once the receivers' windows are known: the keys of receiver r at or
// below its first table entry t[E] (E = kEdgeReserve) or above its last
// t[nb - E - 1] (the model placed the receiver's edge off its splitter)
// would all fall in bucket 0 or the last bucket.  The reserved entries
// split those key ranges evenly by value: entry i < E = lo + (t[E] - lo)
// (i + 1) / (E + 1), lo the receiver's smallest key (the least first key of
// its windows), and likewise from t[nb - E - 1] up to its largest key (the
// ranges are thin slices of the distribution, where the keys lie about
// evenly; a sample of them, one or two keys per sender, is not); then each
// one's position in every sender's window by binary search (bnd index i +
// 1).  One 1024-thread workgroup per receiver: thread (entry, sender).
template <int RW>
__global__ __launch_bounds__(1024) void merge_edges_kernel(const uint32_t* __restrict__ recs, LKey lk,
                                                           const int64_t* __restrict__ seg_begin, int P,
                                                           const int64_t* __restrict__ myC,
                                                           uint64_t* __restrict__ tables,
                                                           int32_t* __restrict__ bnd) {
    constexpr int E = kEdgeReserve;
    __shared__ uint64_t s_lo, s_hi;
    __shared__ uint64_t s_e[2 * E];
    const int r = blockIdx.x, tid = threadIdx.x;
    uint64_t* t = tables + (int64_t)r * kMergeBounds;
    if (tid < 64) {
        uint64_t lo = ~0ull, hi = 0;
        if (tid < P) {
            const int64_t* cj = myC + (int64_t)tid * (P + 1);
            if (cj[r + 1] > cj[r]) {
                uint32_t rr[RW];
                RecIO<RW>::load(recs, seg_begin[tid] + cj[r], rr);
                lo = lkey<RW>(rr, lk);
                RecIO<RW>::load(recs, seg_begin[tid] + cj[r + 1] - 1, rr);
                hi = lkey<RW>(rr, lk);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t x = __shfl_xor(lo, off, 64), y = __shfl_xor(hi, off, 64);
            lo = x < lo ? x : lo;
            hi = y > hi ? y : hi;
        }
        if (tid == 0) {
            s_lo = lo;
            s_hi = hi;
        }
    }
    __syncthreads();
    if (tid < 2 * E) {
        const uint64_t lo = s_lo, hi = s_hi;
        const uint64_t a = t[E], b = t[kMergeBounds - E - 1];
        const int i = tid < E ? tid : tid - E;
        const double f = (double)(i + 1) / (double)(E + 1);
        uint64_t e;
        if (tid < E) {
            // [lo, a] split (a itself when no key lies below it)
            e = lo < a ? lo + (uint64_t)((double)(a - lo) * f) : a;
            e = e > a ? a : e;
        } else {
            e = hi > b ? b + (uint64_t)((double)(hi - b) * f) : b;
            e = e < b ? b : e;
        }
        s_e[tid] = e;
    }
    __syncthreads();
    // non-decreasing (the rounding through double may not keep the order)
    uint64_t ev = 0;
    if (tid < 2 * E) {
        ev = s_e[tid];
        if (tid < E) {
            for (int q = tid + 1; q < E; q++) ev = s_e[q] < ev ? s_e[q] : ev;
        } else {
            for (int q = E; q < tid; q++) ev = s_e[q] > ev ? s_e[q] : ev;
        }
    }
    __syncthreads();
    if (tid < 2 * E) {
        t[tid < E ? tid : kMergeBounds - 2 * E + tid] = ev;
        s_e[tid] = ev;
    }
    __syncthreads();
    // positions: thread (entry q of 2E, sender j)
    for (int w = tid; w < 2 * E * P; w += 1024) {
        const int q = w / P, j = w % P;
        const int k = (q < E ? q : kMergeBounds - 2 * E + q) + 1;
        const uint64_t key = s_e[q];
        const int64_t* cj = myC + (int64_t)j * (P + 1);
        int64_t lo = cj[r], hi = cj[r + 1];
        const int64_t sb = seg_begin[j];
        while (lo < hi) {
            const int64_t m = lo + ((hi - lo) >> 1);
            uint32_t rr[RW];
            RecIO<RW>::load(recs, sb + m, rr);
            if (lkey<RW>(rr, lk) <= key) lo = m + 1; else hi = m;
        }
        bnd[merge_bnd_at(r, j, k, P)] = (int32_t)lo;
    }
}

// The bounds clamped to receiver r's window [myC[j][r], myC[j][r+1]) in
// every sender j (written back: entry k by the thread of (r, k), the last
// entry also by the thread of (r, kMergeBins - 1); an entry read while its
// owner writes it clamps to the same value), and the bucket sizes.  One
// thread per (r, k), walking the senders; the lanes of a wave hold
// neighbouring k, so every read and write is one contiguous run.
__global__ __launch_bounds__(256) void merge_count_kernel(int32_t* __restrict__ bnd, int P,
                                                          const int64_t* __restrict__ myC,
                                                          uint32_t* __restrict__ hist) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (r, k)
    if (g >= (int64_t)P * kMergeBins) return;
    const int k = (int)(g % kMergeBins), r = (int)(g / kMergeBins);
    uint32_t c = 0;
    for (int j = 0; j < P; j++) {
        const int64_t* cj = myC + (int64_t)j * (P + 1);
        const int32_t lo0 = (int32_t)cj[r], hi0 = (int32_t)cj[r + 1];
        int32_t* bk = bnd + merge_bnd_at(r, j, k, P);
        const int32_t a = min(max(bk[0], lo0), hi0), b = min(max(bk[1], lo0), hi0);
        c += (uint32_t)(b - a);
        bk[0] = a;
        if (k == kMergeBins - 1) bk[1] = b;
    }
    hist[g] = c;
}

// groups of the merged SecondSort: tasks[] (receiver by receiver, in bucket
// order within a receiver; receivers in any order), *ntasks entries.  One
// 64-thread workgroup per receiver, one thread per chunk.
__global__ __launch_bounds__(64) void merge_group_kernel(const uint32_t* __restrict__ hist, uint32_t* __restrict__ tasks,
                                                         uint32_t* __restrict__ ntasks) {
    constexpr int NCH = kMergeBins / kGroupChunk;  // chunks per receiver
    static_assert(NCH <= 64 && kMergeBins % kGroupChunk == 0, "one wave walks the chunks");
    __shared__ uint32_t s_cnt[kMergeBins];
    const int r = blockIdx.x, lane = threadIdx.x;
    const uint32_t* h = hist + (int64_t)r * kMergeBins;
    for (int k = lane; k < kMergeBins; k += 64) s_cnt[k] = h[k];
    __syncthreads();
    // the greedy groups of chunk `lane` (counted when out == nullptr, else
    // written): buckets above kMergeCap alone to the second list
    auto walk = [&](uint32_t* out, uint32_t* out_big, uint32_t& nbig) -> uint32_t {
        uint32_t n = 0;
        nbig = 0;
        if (lane < NCH) {
            const int end = (lane + 1) * kGroupChunk;
            for (int k = lane * kGroupChunk; k < end;) {
                const uint32_t code0 = (uint32_t)(((int64_t)r * kMergeBins + k) << 3);
                uint32_t sum = s_cnt[k];
                if (sum > (uint32_t)kMergeCap) {
                    if (out_big) out_big[nbig] = code0;
                    nbig++;
                    k++;
                    continue;
                }
                int len = 1;
                while (k + len < end && len < kGroupMax && sum + s_cnt[k + len] <= (uint32_t)kMergeCap)
                    sum += s_cnt[k + len++];
                if (out) out[n] = code0 | (uint32_t)(len - 1);
                n++;
                k += len;
            }
        }
        return n;
    };
    uint32_t nb = 0;
    const uint32_t n = walk(nullptr, nullptr, nb);
    uint32_t x = n, xb = nb;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64), yb = __shfl_up(xb, off, 64);
        if (lane >= off) {
            x += y;
            xb += yb;
        }
    }
    uint32_t base = 0, base_b = 0;
    if (lane == 63) {
        base = atomicAdd(&ntasks[0], x);
        base_b = atomicAdd(&ntasks[1], xb);
    }
    base = __shfl(base, 63, 64);
    base_b = __shfl(base_b, 63, 64);
    uint32_t nb2 = 0;
    walk(tasks + base + (x - n), tasks + (int64_t)gridDim.x * kMergeBins + base_b + (xb - nb), nb2);
}

// sub-runs of receiver r's buckets [ka, kb) in sender order: s_beg[j] =
// first record (absolute), s_pref[j] = position of sender j's records in the
// bucket, s_pref[P] = the bucket's size (wave 0 writes, the caller syncs)
__device__ __forceinline__ void merge_subruns(const int32_t* __restrict__ bnd, const int64_t* __restrict__ seg_begin,
                                              int P, int r, int ka, int kb, int64_t* s_beg, int* s_pref) {
    const int lane = threadIdx.x & 63;
    int len = 0;
    int64_t beg = 0;
    if (lane < P) {
        const int32_t a = bnd[merge_bnd_at(r, lane, ka, P)];
        beg = seg_begin[lane] + a;
        len = bnd[merge_bnd_at(r, lane, kb, P)] - a;
    }
    int x = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane < P) {
        s_beg[lane] = beg;
        s_pref[lane] = x - len;
    }
    if (lane == P - 1) s_pref[P] = x;
    if (lane > P) s_pref[lane] = 0x7fffffff;  // above every position
}

// The groups' descriptors, one wave per group: what a sorting workgroup
// needs before its first record load, at addresses given by its index
__global__ __launch_bounds__(256) void merge_desc_kernel(const uint32_t* __restrict__ tasks, uint32_t ntasks,
                                                         const int32_t* __restrict__ bnd,
                                                         const int64_t* __restrict__ seg_begin, int P,
                                                         const int64_t* __restrict__ bucket_base,
                                                         const uint64_t* __restrict__ tables,
                                                         MergeTask* __restrict__ desc, int64_t* __restrict__ sub_beg,
                                                         int32_t* __restrict__ sub_pref) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= (int64_t)ntasks) return;
    const uint32_t task = tasks[q];
    const int64_t b = task >> 3;
    const int r = (int)(b / kMergeBins), ka = (int)(b % kMergeBins), kb = ka + (int)(task & 7u) + 1;
    int len = 0;
    int64_t beg = 0;
    if (lane < P) {
        const int32_t a = bnd[merge_bnd_at(r, lane, ka, P)];
        beg = seg_begin[lane] + a;
        len = bnd[merge_bnd_at(r, lane, kb, P)] - a;
    }
    int x = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int total = __shfl(x, 63, 64);
    sub_beg[q * 64 + lane] = beg;
    // positions of the senders' records in the group; entries from P on lie
    // above every position
    sub_pref[q * 64 + lane] = lane < P ? x - len : (lane == P ? total : 0x7fffffff);
    if (lane == 0) {
        const uint64_t* t = tables + (int64_t)r * kMergeBounds;
        MergeTask m;
        m.base = bucket_base[b];
        m.count = (uint32_t)total;
        m.task = task;
        m.tlo = ka > 0 ? t[ka - 1] : 0;
        m.thi = kb - 1 < kMergeBounds ? t[kb - 1] : ~0ull;
        desc[q] = m;
    }
}

// One workgroup per group: task = first bucket << 3 | (buckets - 1).
template <int RW, int ITEMS>
__global__ __launch_bounds__(kMergeBlock, ITEMS == kMergeItems ? 6 : 4) void merge_bucket_kernel(
    const uint32_t* __restrict__ recs, uint32_t* __restrict__ dst, LKey lk,
    const int64_t* __restrict__ seg_begin, const int32_t* __restrict__ bnd, int P,
    const uint32_t* __restrict__ hist, const int64_t* __restrict__ bucket_base,
    const MergeTask* __restrict__ desc, const int64_t* __restrict__ sub_beg, const int32_t* __restrict__ sub_pref,
    int64_t* __restrict__ overflow, uint32_t* __restrict__ noverflow, int max_overflow, uint32_t* __restrict__ hard,
    uint32_t* __restrict__ nhard, uint32_t* __restrict__ medium, uint32_t* __restrict__ nmedium,
    long long* __restrict__ checksum_out, uint32_t ovf_cap) {
    using C = BSort<kMergeBlock, ITEMS>;
    __shared__ int64_t s_beg[kMergeMaxP];
    __shared__ int s_pref[kMergeMaxP + 1];
    __shared__ long long s_sum2[C::W];
    __shared__ uint32_t s_bins[C::BW];
    const int tid = threadIdx.x, wave = tid >> 6;
    const MergeTask m = desc[blockIdx.x];
    if (wave == 0) {
        s_beg[tid] = sub_beg[(int64_t)blockIdx.x * 64 + tid];
        s_pref[tid] = sub_pref[(int64_t)blockIdx.x * 64 + tid];
    }
    const int64_t b = m.task >> 3;  // the group's first bucket
    const int len = (int)(m.task & 7u) + 1;
    const int r = (int)(b / kMergeBins), ka = (int)(b % kMergeBins), kb = ka + len;  // buckets [ka, kb)
    const int count = (int)m.count;
    const int64_t base = m.base;
    const uint64_t tlo = m.tlo, thi = m.thi;
    __syncthreads();
    if (count == 0) return;
    const GatherSrc<RW> gs{recs, s_beg, s_pref, P};
    // a single bucket that holds one key value (see single_bucket)
    const bool single = len == 1 && (ka == 0 ? thi == 0 : (ka < kMergeBounds && tlo + 1 == thi));
    // the keys of buckets [ka, kb) lie in (t[ka-1], t[kb-1]]
    const bool bounded = ka > 0 && kb < kMergeBins && tlo < thi;
    int rc = 1;
    if (!single && count <= C::CAP && (uint32_t)count <= ovf_cap) {
        rc = sort_bucket_lds<RW, kMergeBlock, GatherSrc<RW>, ITEMS>(gs, dst, base, count, lk, s_bins, checksum_out,
                                                                     bounded, bounded ? tlo + 1 : 0,
                                                                     bounded ? thi : ~0ull);
        if (rc == 0) return;
    }
    // copy in sender order, bucket by bucket: final for a single-value
    // bucket, otherwise the input of the radix path
    long long csum = 0;
    for (int q = 0; q < len; q++) {
        const int64_t bq = b + q;
        const int cq = (int)hist[bq];
        if (cq == 0) continue;
        const int64_t baseq = bucket_base[bq];
        if (len > 1) {
            __syncthreads();  // the previous bucket's gathers are done
            if (wave == 0) merge_subruns(bnd, seg_begin, P, r, ka + q, ka + q + 1, s_beg, s_pref);
            __syncthreads();
        }
        for (int i = tid; i < cq; i += C::B) {
            uint32_t x[RW];
            gs.load(i, x);
            RecIO<RW>::store(dst, baseq + i, x);
            if (checksum_out) {
#pragma unroll
                for (int w = 0; w < RW; w++) csum += sbyte_sum(x[w]);
            }
        }
        if (!single && tid == 0) {
            if ((uint32_t)cq > ovf_cap) {
                const uint32_t slot = atomicAdd(noverflow, 1u);  // above the capacity given
                if ((int)slot < max_overflow) {
                    overflow[2 * slot] = baseq;
                    overflow[2 * slot + 1] = (int64_t)cq;
                }
            } else if (cq <= C::CAP && rc == 1) {
                hard[atomicAdd(nhard, 1u)] = (uint32_t)bq;  // a crowded bin: the LDS radix sort
            } else if (cq <= kChunkCap) {
                medium[atomicAdd(nmedium, 1u)] = (uint32_t)bq;  // the 1024-thread in-LDS sort
            } else {
                const uint32_t slot = atomicAdd(noverflow, 1u);
                if ((int)slot < max_overflow) {
                    overflow[2 * slot] = baseq;
                    overflow[2 * slot + 1] = (int64_t)cq;
                }
            }
        }
    }
    if (checksum_out) block_checksum_add<C::W>(csum, s_sum2, checksum_out);
}

// The listed buckets: stable LSD radix sort in LDS over the bytes of
// (key - min) that vary.  Persistent blocks walk the list.
template <int RW>
__global__ __launch_bounds__(kCBlock) void lsd_sort_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, LKey lk,
    const uint32_t* __restrict__ hist, const int64_t* __restrict__ bucket_base,
    const uint32_t* __restrict__ hard, const uint32_t* __restrict__ nhard, long long* __restrict__ checksum_out) {
    __shared__ uint64_t s_key[kChunkCap];
    __shared__ uint16_t s_idx[2][kChunkCap];
    __shared__ uint16_t s_whist[kCWaves][256];
    __shared__ uint32_t s_excl[256];
    __shared__ uint64_t s_red[2 * kCWaves];
    __shared__ uint32_t s_tmp[4];
    __shared__ long long s_sum[kCWaves];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t nh = *nhard;
    long long csum = 0;
    for (uint32_t q = blockIdx.x; q < nh; q += gridDim.x) {
        const int64_t b = hard[q];
        const int64_t base = bucket_base[b];
        const int count = (int)hist[b];
        if (count <= 0 || count > kChunkCap) continue;  // uniform over the block
        uint32_t r[kCItems][RW];
        uint64_t kk[kCItems];
        uint64_t mn, mx;
        __syncthreads();  // LDS of the previous bucket is free
        load_bucket<RW>(src, base, count, lk, r, kk, s_red, mn, mx);
        const uint64_t span = mx - mn;
        // at least one pass: it writes the order the rank step reads, also
        // when every key is equal
        const int nbytes = max(1, (64 - __clzll((long long)span) + 7) >> 3);
#pragma unroll
        for (int k = 0; k < kCItems; k++) {
            const int i = wave * kCWspan + k * 64 + lane;
            if (i < count) s_key[i] = kk[k] - mn;
        }
        __syncthreads();
        int cur = 0;
        for (int p = 0; p < nbytes; p++) {
            const int sh = 8 * p;
            lds_count_pass(p == 0 ? nullptr : s_idx[cur], s_idx[cur ^ (p == 0 ? 0 : 1)], count,
                           [&](uint16_t e) { return (uint32_t)((s_key[e] >> sh) & 0xffu); }, s_whist, s_excl,
                           s_tmp);
            cur ^= (p == 0 ? 0 : 1);
        }
        // rank of every original record, then a direct write
        for (int o = tid; o < count; o += kCBlock) s_idx[cur ^ 1][s_idx[cur][o]] = (uint16_t)o;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kCItems; k++) {
            const int i = wave * kCWspan + k * 64 + lane;
            if (i < count) {
                RecIO<RW>::store(dst, base + s_idx[cur ^ 1][i], r[k]);
                if (checksum_out) {
#pragma unroll
                    for (int w = 0; w < RW; w++) csum += sbyte_sum(r[k][w]);
                }
            }
        }
    }
    if (checksum_out) block_checksum_add(csum, s_sum, checksum_out);
}

// The listed segments ({begin, count} pairs) that hold one key value are in
// stable order already.  one_value_check_kernel: gridDim.x workgroups per
// segment (blockIdx.y) compare their share of its keys with its first key;
// a segment with another key gets the top bit of its count set (negative).
// one_value_copy_kernel: the segments left non-negative are copied to dst
// (when it is another buffer) and added to the checksum.
template <int RW>
__global__ __launch_bounds__(kBlock) void one_value_check_kernel(const uint32_t* __restrict__ src, LKey lk,
                                                                 int64_t* __restrict__ list) {
    const int64_t base = list[2 * (int64_t)blockIdx.y], cnt = list[2 * (int64_t)blockIdx.y + 1];
    if (cnt <= 0) return;  // empty, or another workgroup found a second key value
    uint32_t r0[RW];
    RecIO<RW>::load(src, base, r0);
    const uint64_t k0 = lkey<RW>(r0, lk);
    bool diff = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock) {
        uint32_t rr[RW];
        RecIO<RW>::load(src, base + i, rr);
        diff = diff || lkey<RW>(rr, lk) != k0;
    }
    if (__ballot(diff) != 0 && (threadIdx.x & 63) == 0)
        atomicOr(reinterpret_cast<unsigned int*>(&list[2 * (int64_t)blockIdx.y + 1]) + 1, 0x80000000u);
}

template <int RW>
__global__ __launch_bounds__(kBlock) void one_value_copy_kernel(const uint32_t* __restrict__ src,
                                                                uint32_t* __restrict__ dst,
                                                                const int64_t* __restrict__ list,
                                                                long long* __restrict__ checksum_out) {
    __shared__ long long s_sum[kWaves];
    const int64_t base = list[2 * (int64_t)blockIdx.y], cnt = list[2 * (int64_t)blockIdx.y + 1];
    if (cnt <= 0) return;
    long long csum = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock) {
        uint32_t rr[RW];
        RecIO<RW>::load(src, base + i, rr);
        if (src != dst) RecIO<RW>::store(dst, base + i, rr);
#pragma unroll
        for (int w = 0; w < RW; w++) csum += sbyte_sum(rr[w]);
    }
    if (checksum_out) block_checksum_add<kWaves>(csum, s_sum, checksum_out);
}

// range[j][b] = first entry above segment j's FirstSort boundary b - 1 (the
// entries of bucket b are range[j][b] .. range[j][b + 1] - 1).  One thread
// per (j, b), b = 0 .. kLBins.
__global__ void entry_ranges_kernel(const uint64_t* __restrict__ ftab, int64_t ftstride, int P,
                                    const uint64_t* __restrict__ E, uint32_t* __restrict__ range) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)P * (kLBins + 1)) return;
    const int j = (int)(g / (kLBins + 1)), b = (int)(g % (kLBins + 1));
    const uint32_t ne = (uint32_t)P * kMergeBounds;
    uint32_t res;
    if (b == 0) {
        res = 0;
    } else if (b == kLBins) {
        res = ne;
    } else {
        const uint64_t t = ftab[(int64_t)j * ftstride + b - 1];
        uint32_t a = 0, e = ne;
        while (a < e) {
            const uint32_t m = (a + e) >> 1;
            if (E[m] <= t) a = m + 1; else e = m;
        }
        res = a;
    }
    range[g] = res;
}

// The entries of every FirstSort bucket not marked done (sorted by the LDS
// radix sort, the overflow fallback or a plain sort of the whole side): a
// binary search per entry in the sorted bucket.  One workgroup per bucket;
// the last P * P workgroups write the fixed bounds k = 0 and k = kMergeBins.
template <int RW>
__global__ __launch_bounds__(kBlock) void entries_fixup_kernel(const uint32_t* __restrict__ recs, LKey lk,
                                                               EntryOut eo, const uint32_t* __restrict__ hist,
                                                               const int64_t* __restrict__ bucket_base,
                                                               const int64_t* __restrict__ seg_count) {
    const int64_t nb = (int64_t)eo.P * kLBins;
    const int64_t bb = blockIdx.x;
    if (bb >= nb) {
        const int64_t q = bb - nb;  // (r, j)
        const int r = (int)(q / eo.P), j = (int)(q % eo.P);
        if (threadIdx.x == 0) {
            eo.bnd[merge_bnd_at(r, j, 0, eo.P)] = 0;
            eo.bnd[merge_bnd_at(r, j, kMergeBins, eo.P)] = (int32_t)seg_count[j];
        }
        return;
    }
    if (eo.done[bb]) return;
    const int64_t base = bucket_base[bb];
    const int64_t cnt = hist[bb];
    write_entries<kBlock>(eo, (int)(bb / kLBins), (int)(bb % kLBins), base, [&](uint64_t t) -> uint32_t {
        int64_t a = 0, e = cnt;
        while (a < e) {
            const int64_t m = (a + e) >> 1;
            uint32_t rr[RW];
            RecIO<RW>::load(recs, base + m, rr);
            if (lkey<RW>(rr, lk) <= t) a = m + 1; else e = m;
        }
        return (uint32_t)a;
    });
}

// ---------------------------------------------------------------------------
// host launchers
int launch_lhist(int rw, const void* src, const SegDev& sd, int64_t ntiles, int tile_recs, const LKey& lk,
                 const uint64_t* tables, int64_t tstride, uint32_t* hist, uint16_t* tile_cnt, long long* checksum,
                 uint16_t* bkt, hipStream_t st) {
    if (ntiles <= 0) return 0;
    // one round of resident workgroups: each takes an equal run of tiles
    const int64_t slots = rw == 2 ? resident_blocks(lhist_kernel<2>, kBlock) : resident_blocks(lhist_kernel<4>, kBlock);
    const int64_t tpb = (ntiles + slots - 1) / slots;
    const int64_t blocks = (ntiles + tpb - 1) / tpb;
    const uint32_t* s = (const uint32_t*)src;
    switch (rw) {
        case 2: hipLaunchKernelGGL(lhist_kernel<2>, dim3((unsigned)blocks), dim3(kBlock), 0, st, s, sd, ntiles, tpb, tile_recs, lk, tables, tstride, hist, tile_cnt, checksum, bkt); break;
        case 4: hipLaunchKernelGGL(lhist_kernel<4>, dim3((unsigned)blocks), dim3(kBlock), 0, st, s, sd, ntiles, tpb, tile_recs, lk, tables, tstride, hist, tile_cnt, checksum, bkt); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_lscan(int nseg, const uint32_t* hist, const int64_t* seg_begin, int64_t* base, hipStream_t st, int nb) {
    if (nb % 1024 != 0 || nb > 4096) return -1;
    hipLaunchKernelGGL(lscan_kernel, dim3(nseg), dim3(1024), 0, st, hist, seg_begin, base, nb);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int lscatter_tile(int rw) { return rw == 2 ? 1024 * 8 : 512 * 6; }

int launch_ltile_scan(const uint16_t* tile_cnt, const SegDev& sd, const int64_t* bucket_base, uint32_t* tile_off,
                      hipStream_t st) {
    if (sd.nseg <= 0) return 0;
    hipLaunchKernelGGL(ltile_scan_kernel, dim3((unsigned)sd.nseg * (kLBins / 256)), dim3(256), 0, st, tile_cnt, sd,
                       bucket_base, tile_off);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_lscatter(int rw, const void* src, void* dst, void* eq_dst, const SegDev& sd, int64_t ntiles,
                    const uint16_t* bkt, const uint64_t* tables, int64_t tstride, const uint32_t* tile_off,
                    uint32_t* err, hipStream_t st) {
    if (ntiles <= 0) return 0;
    const unsigned grid = (unsigned)(8 * ((ntiles + 7) / 8));
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    uint32_t* e = (uint32_t*)eq_dst;
    switch (rw) {
        case 2: hipLaunchKernelGGL((lscatter_kernel<2, 8, 1024>), dim3(grid), dim3(1024), 0, st, s, d, sd, ntiles, bkt, tables, tstride, tile_off, err, e); break;
        case 4: hipLaunchKernelGGL((lscatter_kernel<4, 6, 512>), dim3(grid), dim3(512), 0, st, s, d, sd, ntiles, bkt, tables, tstride, tile_off, err, e); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_chunk_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                      const int64_t* bucket_base, const uint64_t* tables, int64_t tstride, int64_t nbuckets,
                      int64_t* overflow, uint32_t* noverflow, int max_overflow, uint32_t* hard, uint32_t* nhard,
                      long long* checksum, bool eq_done, hipStream_t st, hipEvent_t mid, uint32_t ovf_cap) {
    if (nbuckets <= 0) {
        if (mid) MPSG_CHECK(hipEventRecord(mid, st));
        return 0;
    }
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    const int ed = eq_done ? 1 : 0;
    const unsigned persistent = (unsigned)std::min<int64_t>(nbuckets, 512);
    const unsigned grid = (unsigned)nbuckets;
    switch (rw) {
        case 2:
            hipLaunchKernelGGL((interp_sort_kernel<2, kCBlock>), dim3(grid), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base,
                               tables, tstride, nbuckets, overflow, noverflow, max_overflow, hard, nhard, nullptr, nullptr,
                               checksum, ed, EntryOut{}, ovf_cap);
            MPSG_CHECK(hipGetLastError());
            if (mid) MPSG_CHECK(hipEventRecord(mid, st));
            hipLaunchKernelGGL(lsd_sort_kernel<2>, dim3(persistent), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base,
                               (const uint32_t*)hard, (const uint32_t*)nhard, checksum);
            break;
        case 4:
            hipLaunchKernelGGL((interp_sort_kernel<4, kCBlock>), dim3(grid), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base,
                               tables, tstride, nbuckets, overflow, noverflow, max_overflow, hard, nhard, nullptr, nullptr,
                               checksum, ed, EntryOut{}, ovf_cap);
            MPSG_CHECK(hipGetLastError());
            if (mid) MPSG_CHECK(hipEventRecord(mid, st));
            hipLaunchKernelGGL(lsd_sort_kernel<4>, dim3(persistent), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base,
                               (const uint32_t*)hard, (const uint32_t*)nhard, checksum);
            break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_segments_lds(int rw, int form, const void* src, void* dst, const LKey& lk, const uint32_t* cnt,
                        const int64_t* base, int64_t na, int64_t nb, int64_t* overflow, uint32_t* noverflow,
                        uint32_t* hard, uint32_t* nhard_a, uint32_t* nhard_b, hipStream_t st) {
    if (form == 0 || na == 0)
        return launch_chunk_sort(rw, src, dst, lk, cnt, base, nullptr, 0, na + nb, overflow, noverflow, 1, hard,
                                 nhard_a, nullptr, false, st);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    const unsigned grid = (unsigned)na, persistent = (unsigned)std::min<int64_t>(na, 512);
    const uint32_t cap = (uint32_t)kChunkCap;
#define MPSG_SEGFORM(R, I) do { \
        hipLaunchKernelGGL((interp_sort_kernel<R, kMergeBlock, I>), dim3(grid), dim3(kMergeBlock), 0, st, s, d, lk, \
                           cnt, base, nullptr, 0, na, overflow, noverflow, 1, hard, nhard_a, nullptr, nullptr, nullptr, \
                           0, EntryOut{}, cap); \
        MPSG_CHECK(hipGetLastError()); \
        hipLaunchKernelGGL(lsd_sort_kernel<R>, dim3(persistent), dim3(kCBlock), 0, st, s, d, lk, cnt, base, \
                           (const uint32_t*)hard, (const uint32_t*)nhard_a, nullptr); \
    } while (0)
    switch (rw * 4 + form) {
        case 9: MPSG_SEGFORM(2, kSmallItems); break;
        case 10: MPSG_SEGFORM(2, kMergeItems); break;
        case 17: MPSG_SEGFORM(4, kSmallItems); break;
        case 18: MPSG_SEGFORM(4, kMergeItems); break;
        default: return -1;
    }
#undef MPSG_SEGFORM
    MPSG_CHECK(hipGetLastError());
    if (nb == 0) return 0;
    // the segments above the form's capacity: the 1024-thread form
    return launch_chunk_sort(rw, src, dst, lk, cnt + na, base + na, nullptr, 0, nb, overflow, noverflow, 1, hard + na,
                             nhard_b, nullptr, false, st);
}

int segments_lds_cap(int form) {
    return form == 1 ? BSort<kMergeBlock, kSmallItems>::CAP : (form == 2 ? BSort<kMergeBlock, kMergeItems>::CAP : kChunkCap);
}

int launch_chunk_sort_small(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                            const int64_t* bucket_base, const uint64_t* tables, int64_t tstride, int64_t nbuckets,
                            int64_t* overflow, uint32_t* noverflow, int max_overflow, uint32_t* hard, uint32_t* nhard,
                            uint32_t* medium, uint32_t* nmedium, long long* checksum, bool eq_done, hipStream_t st,
                            const EntryOut* eop, uint32_t ovf_cap) {
    if (nbuckets <= 0) return 0;
    const EntryOut eo = eop ? *eop : EntryOut{};
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    const int ed = eq_done ? 1 : 0;
    const unsigned grid = (unsigned)nbuckets;
    switch (rw) {
        case 2:
            hipLaunchKernelGGL((interp_sort_kernel<2, kMergeBlock, kSmallItems>), dim3(grid), dim3(kMergeBlock), 0, st, s, d, lk, hist,
                               bucket_base, tables, tstride, nbuckets, overflow, noverflow, max_overflow, hard, nhard,
                               medium, nmedium, checksum, ed, eo, ovf_cap);
            break;
        case 4:
            hipLaunchKernelGGL((interp_sort_kernel<4, kMergeBlock, kSmallItems>), dim3(grid), dim3(kMergeBlock), 0, st, s, d, lk, hist,
                               bucket_base, tables, tstride, nbuckets, overflow, noverflow, max_overflow, hard, nhard,
                               medium, nmedium, checksum, ed, eo, ovf_cap);
            break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_bounds(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, const int64_t* seg_count,
                        int P, const uint64_t* tables, const uint64_t* ftab, int64_t ftstride, const int64_t* fbase,
                        int32_t* bnd, hipStream_t st, const int64_t* win, const uint64_t* samp,
                        const int64_t* spref) {
    const int64_t n = (int64_t)P * P * (kMergeBins + 1);
    const unsigned g = (unsigned)((n + 255) / 256);
    const uint32_t* r = (const uint32_t*)recs;
    switch (rw) {
        case 2: hipLaunchKernelGGL(merge_bounds_kernel<2>, dim3(g), dim3(256), 0, st, r, lk, seg_begin, seg_count, P, tables, ftab, ftstride, fbase, win, samp, spref, bnd); break;
        case 4: hipLaunchKernelGGL(merge_bounds_kernel<4>, dim3(g), dim3(256), 0, st, r, lk, seg_begin, seg_count, P, tables, ftab, ftstride, fbase, win, samp, spref, bnd); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_sample_tables(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, int P,
                               const int64_t* myC, uint64_t* tables, uint64_t* samp, int64_t* spref,
                               hipStream_t st) {
    if (P < 1 || P > kMergeMaxP) return -1;
    const uint32_t* r = (const uint32_t*)recs;
    switch (rw) {
        case 2: hipLaunchKernelGGL(merge_sample_tables_kernel<2>, dim3(P), dim3(1024), 0, st, r, lk, seg_begin, P, myC, tables, samp, spref); break;
        case 4: hipLaunchKernelGGL(merge_sample_tables_kernel<4>, dim3(P), dim3(1024), 0, st, r, lk, seg_begin, P, myC, tables, samp, spref); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_seg_sample(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, const int64_t* seg_count,
                      int P, int S, uint64_t* out, hipStream_t st) {
    const int64_t n = (int64_t)P * S;
    const unsigned g = (unsigned)((n + 255) / 256);
    const uint32_t* r = (const uint32_t*)recs;
    switch (rw) {
        case 2: hipLaunchKernelGGL(seg_sample_kernel<2>, dim3(g), dim3(256), 0, st, r, lk, seg_begin, seg_count, P, S, out); break;
        case 4: hipLaunchKernelGGL(seg_sample_kernel<4>, dim3(g), dim3(256), 0, st, r, lk, seg_begin, seg_count, P, S, out); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_sample_tables(const uint64_t* sorted, int P, int S, uint64_t* tables, hipStream_t st) {
    if (S < kLBins || S % kLBins) return -1;
    const int64_t n = (int64_t)P * kLBounds;
    hipLaunchKernelGGL(sample_tables_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sorted, P, S, tables);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_tables_reserve_edges(uint64_t* tables, int P, hipStream_t st) {
    const int n = P * 2 * kEdgeReserve;
    hipLaunchKernelGGL(tables_reserve_edges_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tables, P);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_edges(int rw, const void* recs, const LKey& lk, const int64_t* seg_begin, int P, const int64_t* myC,
                       uint64_t* tables, int32_t* bnd, hipStream_t st) {
    if (P < 1 || P > kMergeMaxP) return -1;
    const uint32_t* r = (const uint32_t*)recs;
    switch (rw) {
        case 2: hipLaunchKernelGGL(merge_edges_kernel<2>, dim3(P), dim3(1024), 0, st, r, lk, seg_begin, P, myC, tables, bnd); break;
        case 4: hipLaunchKernelGGL(merge_edges_kernel<4>, dim3(P), dim3(1024), 0, st, r, lk, seg_begin, P, myC, tables, bnd); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_count(int32_t* bnd, int P, const int64_t* myC, uint32_t* hist, hipStream_t st) {
    const int64_t n = (int64_t)P * kMergeBins;  // one thread each
    hipLaunchKernelGGL(merge_count_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, bnd, P, myC, hist);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_group(const uint32_t* hist, int P, uint32_t* tasks, uint32_t* ntasks, hipStream_t st) {
    if (P <= 0) return 0;
    hipLaunchKernelGGL(merge_group_kernel, dim3((unsigned)P), dim3(64), 0, st, hist, tasks, ntasks);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_desc(const uint32_t* tasks, uint32_t ntasks, const int32_t* bnd, const int64_t* seg_begin, int P,
                      const int64_t* bucket_base, const uint64_t* tables, MergeTask* desc, int64_t* sub_beg,
                      int32_t* sub_pref, hipStream_t st) {
    if (ntasks == 0) return 0;
    if (P > kMergeMaxP) return -1;
    hipLaunchKernelGGL(merge_desc_kernel, dim3((ntasks + 3) / 4), dim3(256), 0, st, tasks, ntasks, bnd, seg_begin, P,
                       bucket_base, tables, desc, sub_beg, sub_pref);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_merge_bucket(int rw, bool big, const void* recs, void* dst, const LKey& lk, const int64_t* seg_begin,
                        const int32_t* bnd, int P, const uint32_t* hist, const int64_t* bucket_base,
                        const MergeTask* desc, const int64_t* sub_beg,
                        const int32_t* sub_pref, uint32_t ntasks, int64_t* overflow, uint32_t* noverflow,
                        int max_overflow, uint32_t* hard, uint32_t* nhard, uint32_t* medium, uint32_t* nmedium,
                        long long* checksum, hipStream_t st, uint32_t ovf_cap) {
    if (P > kMergeMaxP) return -1;
    if (ntasks == 0) return 0;
    const uint32_t* r = (const uint32_t*)recs;
    uint32_t* d = (uint32_t*)dst;
#define MPSG_MERGE(R, I) hipLaunchKernelGGL((merge_bucket_kernel<R, I>), dim3(ntasks), dim3(kMergeBlock), 0, st, r, d, lk, seg_begin, bnd, P, hist, bucket_base, desc, sub_beg, sub_pref, overflow, noverflow, max_overflow, hard, nhard, medium, nmedium, checksum, ovf_cap)
    switch (rw * 2 + (big ? 1 : 0)) {
        case 4: MPSG_MERGE(2, kMergeItems); break;
        case 5: MPSG_MERGE(2, kSmallItems); break;
        case 8: MPSG_MERGE(4, kMergeItems); break;
        case 9: MPSG_MERGE(4, kSmallItems); break;
        default: return -1;
    }
#undef MPSG_MERGE
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_listed_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* list, uint32_t nlist,
                       const uint32_t* hist, const int64_t* bucket_base, const uint64_t* tables, int nbins,
                       int64_t tstride, uint32_t* hard, uint32_t* nhard, long long* checksum, hipStream_t st,
                       const EntryOut* eop) {
    if (nlist == 0) return 0;
    const EntryOut eo = eop ? *eop : EntryOut{};
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    switch (rw) {
        case 2: hipLaunchKernelGGL(listed_sort_kernel<2>, dim3(nlist), dim3(kCBlock), 0, st, s, d, lk, list, hist, bucket_base, tables, nbins, tstride, hard, nhard, checksum, eo); break;
        case 4: hipLaunchKernelGGL(listed_sort_kernel<4>, dim3(nlist), dim3(kCBlock), 0, st, s, d, lk, list, hist, bucket_base, tables, nbins, tstride, hard, nhard, checksum, eo); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_lsd_sort(int rw, const void* src, void* dst, const LKey& lk, const uint32_t* hist,
                    const int64_t* bucket_base, const uint32_t* hard, const uint32_t* nhard, int64_t max_buckets,
                    long long* checksum, hipStream_t st) {
    if (max_buckets <= 0) return 0;
    const unsigned persistent = (unsigned)std::min<int64_t>(max_buckets, 512);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    switch (rw) {
        case 2: hipLaunchKernelGGL(lsd_sort_kernel<2>, dim3(persistent), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base, hard, nhard, checksum); break;
        case 4: hipLaunchKernelGGL(lsd_sort_kernel<4>, dim3(persistent), dim3(kCBlock), 0, st, s, d, lk, hist, bucket_base, hard, nhard, checksum); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_one_value_filter(int rw, const void* src, void* dst, const LKey& lk, int64_t* list, uint32_t nlist,
                            int64_t max_count, long long* checksum, hipStream_t st) {
    if (nlist == 0) return 0;
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    // about 16 records per thread of the largest segment, at most 1024
    // workgroups per segment
    const int64_t per = (int64_t)kBlock * 16;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (max_count + per - 1) / per));
    const dim3 grid(gx, nlist);
    const bool copy = src != dst || checksum != nullptr;
    switch (rw) {
        case 2:
            hipLaunchKernelGGL(one_value_check_kernel<2>, grid, dim3(kBlock), 0, st, s, lk, list);
            if (copy) hipLaunchKernelGGL(one_value_copy_kernel<2>, grid, dim3(kBlock), 0, st, s, d, list, checksum);
            break;
        case 4:
            hipLaunchKernelGGL(one_value_check_kernel<4>, grid, dim3(kBlock), 0, st, s, lk, list);
            if (copy) hipLaunchKernelGGL(one_value_copy_kernel<4>, grid, dim3(kBlock), 0, st, s, d, list, checksum);
            break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_entry_ranges(const uint64_t* ftab, int64_t ftstride, int P, const uint64_t* E, uint32_t* range,
                        hipStream_t st) {
    const int64_t n = (int64_t)P * (kLBins + 1);
    hipLaunchKernelGGL(entry_ranges_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ftab, ftstride, P, E,
                       range);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_entries_fixup(int rw, const void* recs, const LKey& lk, const EntryOut& eo, const uint32_t* hist,
                         const int64_t* bucket_base, const int64_t* seg_count, hipStream_t st) {
    const unsigned grid = (unsigned)((int64_t)eo.P * kLBins + (int64_t)eo.P * eo.P);
    const uint32_t* r = (const uint32_t*)recs;
    switch (rw) {
        case 2: hipLaunchKernelGGL(entries_fixup_kernel<2>, dim3(grid), dim3(kBlock), 0, st, r, lk, eo, hist, bucket_base, seg_count); break;
        case 4: hipLaunchKernelGGL(entries_fixup_kernel<4>, dim3(grid), dim3(kBlock), 0, st, r, lk, eo, hist, bucket_base, seg_count); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

}  // namespace mpsg
