// onesweep.hip -- stable segmented LSD radix sort of sort records (8-bit
// digits), the device form of the local sorts of the histogram sort
// (FirstSort mpsort-mpi.c:369 and SecondSort mpsort-mpi.c:597, which call
// radix_sort radixsort.c:35-44, a stable merge sort, stdlib/msort.c:52-174).
//
// Per sort:  one upfront pass counts every digit place of every segment
// (hist_kernel), a scan turns the counts into per-segment digit bases and
// marks digit places that hold a single value in every segment (such a pass
// would be the identity permutation and is skipped), then one onesweep pass
// per remaining digit place moves the records:
//   * tiles of BLOCK threads x ITEMS records, tile ids from an atomic ticket so
//     every predecessor tile is already resident (forward progress for the
//     look-back);
//   * wave-striped loads (item k of lane l is record k*64+l of the wave's
//     slice), so ranking items in (k, lane) order is input order: stable;
//   * per-item digit ranking by 8 wave ballots (match-any), per-wave digit
//     counters in LDS;
//   * decoupled look-back over the tiles of the same segment, one thread per
//     digit, 64-bit flag words tagged with a per-launch epoch, kLookback
//     predecessor flags read per step;
//   * records are staged in LDS in digit order and written out in runs that
//     are contiguous per digit (coalesced).
// The tile shape (BLOCK, ITEMS) is chosen per call (TileCfg).
#include <stdlib.h>
#include <stdio.h>

#include "device.h"

namespace mpsg {

// ---------------------------------------------------------------------------
// upfront histogram of all digit places.  Block b counts tiles
// [b*tpb, (b+1)*tpb) in LDS and flushes to the segment's global histogram
// whenever the segment changes, so global atomics scale with blocks, not
// records.
template <int RW>
__global__ __launch_bounds__(kBlock) void hist_kernel(const uint32_t* __restrict__ src, SegDev seg,
                                                      int64_t ntiles, int64_t tpb, int tile_recs,
                                                      int npass, PassTable pt, int b0,
                                                      uint32_t* __restrict__ hist,
                                                      long long* __restrict__ checksum) {
    extern __shared__ uint32_t s_hist[];  // npass * kBins counters (dynamic: short keys leave room
                                          // for more resident workgroups)
    __shared__ long long s_sum[kWaves];
    const int tid = threadIdx.x;
    const int nb = npass * kBins;
    for (int i = tid; i < nb; i += kBlock) s_hist[i] = 0;
    __syncthreads();

    const int64_t t0 = (int64_t)blockIdx.x * tpb;
    const int64_t t1 = min(ntiles, t0 + tpb);
    long long csum = 0;
    int cur = -1;
    // U records per thread at a time; the next U records' loads are issued
    // before this group is counted, so they are in flight meanwhile (the
    // next group may be the next tile's first).  Every thread runs the same
    // number of groups per tile (positions past the tile are ignored).
    constexpr int U = 4, STEP = U * kBlock;
    auto tile_range = [&](int64_t tl, int& s, int64_t& b, int64_t& e) {
        s = upper_index<int64_t>(seg.tile_start, seg.nseg, tl);
        const int64_t sb = seg.begin[s];
        b = sb + (tl - seg.tile_start[s]) * tile_recs;
        e = min(sb + seg.count[s], b + (int64_t)tile_recs);
    };
    auto load_group = [&](uint32_t (&r)[U][RW], int64_t i0, int64_t end) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (i0 + (int64_t)u * kBlock < end) {
                RecIO<RW>::load(src, i0 + (int64_t)u * kBlock, r[u]);
            } else {
#pragma unroll
                for (int w = 0; w < RW; w++) r[u][w] = 0;
            }
        }
    };
    int s_nx = 0;
    int64_t b_nx = 0, e_nx = 0;
    uint32_t r[U][RW];
    if (t0 < t1) tile_range(t0, s_nx, b_nx, e_nx);
    load_group(r, b_nx + tid, e_nx);
    for (int64_t tile = t0; tile < t1; tile++) {
        const int s = s_nx;
        const int64_t begin = b_nx, end = e_nx;
        if (tile + 1 < t1) {
            tile_range(tile + 1, s_nx, b_nx, e_nx);
        } else {
            b_nx = e_nx = 0;
        }
        if (s != cur) {
            if (cur >= 0) {
                __syncthreads();
                uint32_t* h = hist + (int64_t)cur * nb;
                for (int i = tid; i < nb; i += kBlock) {
                    const uint32_t c = s_hist[i];
                    if (c) atomicAdd(&h[i], c);
                    s_hist[i] = 0;
                }
                __syncthreads();
            }
            cur = s;
        }
        const int ngroups = (int)((end - begin + STEP - 1) / STEP);
        for (int q = 0; q < ngroups; q++) {
            const int64_t i0 = begin + tid + (int64_t)q * STEP;
            uint32_t rn[U][RW];
            if (q + 1 < ngroups) {
                load_group(rn, i0 + STEP, end);
            } else {
                load_group(rn, b_nx + tid, e_nx);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (i0 + (int64_t)u * kBlock < end) {
                    if (checksum) {
#pragma unroll
                        for (int w = 0; w < RW; w++) csum += sbyte_sum(r[u][w]);
                    }
                    count_digits<RW>(r[u], npass, pt, b0, s_hist);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int w = 0; w < RW; w++) r[u][w] = rn[u][w];
        }
    }
    __syncthreads();
    if (cur >= 0) {
        uint32_t* h = hist + (int64_t)cur * nb;
        for (int i = tid; i < nb; i += kBlock) {
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

// per (segment, digit place): exclusive scan -> absolute digit bases; flag the
// digit place as non-trivial when the segment holds more than one digit value
__global__ __launch_bounds__(kBlock) void scan_kernel(const uint32_t* __restrict__ hist,
                                                      const int64_t* __restrict__ seg_begin,
                                                      int npass, int64_t* __restrict__ digit_base,
                                                      uint32_t* __restrict__ pass_flags) {
    __shared__ uint32_t s_tmp[kWaves];
    const int64_t sp = blockIdx.x;
    const int s = (int)(sp / npass), p = (int)(sp % npass);
    const int tid = threadIdx.x;
    const uint32_t c = hist[sp * kBins + tid];
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(c, s_tmp, &tot);
    digit_base[sp * kBins + tid] = seg_begin[s] + (int64_t)ex;
    const int nz = __syncthreads_count(c > 0);
    if (tid == 0 && nz > 1) atomicOr(&pass_flags[p], 1u);
}

// exclusive scan of one value per digit held by threads 0..255 of a block of
// BLOCK >= 256 threads; every thread of the block must call it
template <int BLOCK>
__device__ __forceinline__ uint32_t digit_exclusive_scan(uint32_t v, uint32_t* s_tmp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
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
    __syncthreads();
    return pre + x - v;
}

// ---------------------------------------------------------------------------
// one onesweep digit pass
template <int RW, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void onesweep_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, SegDev seg,
    const int64_t* __restrict__ digit_base, int npass, int pass_idx, PassDesc pd,
    uint64_t* __restrict__ lookback, uint32_t* __restrict__ ticket, uint32_t epoch,
    const uint32_t* __restrict__ epoch_base, long long* __restrict__ checksum_out,
    uint32_t* __restrict__ error_flag) {
    constexpr int NW = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WAVE_TILE = 64 * ITEMS;
    __shared__ __attribute__((aligned(16))) uint32_t s_rec[TILE * RW];
    __shared__ uint16_t s_whist[NW][kBins];  // counts and offsets < TILE <= 8192
    __shared__ uint32_t s_excl[kBins];
    __shared__ int64_t s_gdelta[kBins];
    __shared__ int64_t s_tile;
    __shared__ uint32_t s_tmp[4];
    __shared__ long long s_sum[NW];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // replayed launches (graphs) take their epoch from a device counter
    if (epoch_base) epoch += *epoch_base;
    if (tid == 0) s_tile = (int64_t)atomicAdd(ticket, 1u);
    for (int i = tid; i < NW * kBins; i += BLOCK) (&s_whist[0][0])[i] = 0;
    __syncthreads();

    const int64_t tile = s_tile;
    const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
    const int64_t t = tile - seg.tile_start[s];
    const int64_t sbeg = seg.begin[s];
    const int64_t send = sbeg + seg.count[s];
    const int64_t begin = sbeg + t * TILE;
    const int64_t end = min(send, begin + (int64_t)TILE);
    const int count = (int)(end - begin);

    // every load issued before any is used (all ITEMS in flight at once);
    // slots past the tile read its last record again and are not records
    uint32_t r[ITEMS][RW];
    uint32_t dg[ITEMS];
    uint32_t loc[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + wave * WAVE_TILE + k * 64 + lane;
        if (end > begin) RecIO<RW>::load(src, i < end ? i : end - 1, r[k]);
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + wave * WAVE_TILE + k * 64 + lane;
        dg[k] = i < end ? digit_of<RW>(r[k], pd) : (uint32_t)kBins;  // kBins: not a record
    }

    // rank inside the wave: items in (k, lane) order are in input order
    volatile uint16_t* wh = s_whist[wave];
    const uint64_t lt_mask = __lanemask_lt();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t d = dg[k];
        const bool valid = d < kBins;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const uint32_t cnt = __popcll(peers);
        uint32_t base = 0;
        if (valid) base = wh[d];
        loc[k] = base + rank;
        if (valid && rank == 0) wh[d] = (uint16_t)(base + cnt);
    }
    __syncthreads();

    // threads 0..255 own one digit each: wave offsets, tile-local digit start
    const int d = tid;
    uint32_t total = 0;
    if (d < kBins) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t c = s_whist[w][d];
            s_whist[w][d] = (uint16_t)acc;
            acc += c;
        }
        total = acc;
    }
    const uint32_t excl = digit_exclusive_scan<BLOCK>(total, s_tmp);
    if (d < kBins) {
        s_excl[d] = excl;
        int64_t prefix = 0;
        uint64_t* my = lookback + tile * kBins + d;
        if (t == 0) {
            store_agent(my, make_flag(epoch, 1, total));
        } else {
            store_agent(my, make_flag(epoch, 0, total));
            // windowed look-back: read kLookback predecessor flags at once,
            // sum aggregates back to the first inclusive prefix; a flag that
            // is not ready yet is polled again.  The first tile of a segment
            // always publishes an inclusive prefix, so the walk never leaves
            // the segment.
            const int64_t first = seg.tile_start[s];
            int64_t j = tile - 1;
            uint32_t spins = 0;
            while (true) {
                uint64_t f[kLookback];
#pragma unroll
                for (int q = 0; q < kLookback; q++)
                    f[q] = (j - q >= first) ? load_agent(lookback + (j - q) * kBins + d) : 0ull;
                int consumed = 0;
                bool stop = false, done = false;
#pragma unroll
                for (int q = 0; q < kLookback; q++) {
                    if (!stop && j - q >= first) {
                        if ((uint32_t)(f[q] >> 41) != epoch) {
                            stop = true;
                        } else {
                            prefix += (int64_t)(f[q] & kFlagValueMask);
                            consumed = q + 1;
                            if ((f[q] >> 40) & 1ull) { done = true; stop = true; }
                        }
                    }
                }
                if (done) break;
                j -= consumed;
                if (consumed == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(error_flag, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            store_agent(my, make_flag(epoch, 1, (uint64_t)(prefix + total)));
        }
        s_gdelta[d] = digit_base[((int64_t)s * npass + pass_idx) * kBins + d] + prefix - (int64_t)excl;
    }
    __syncthreads();

    // stage records in LDS in (digit, input) order
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t dd = dg[k];
        if (dd < kBins) {
            const uint32_t pos = s_excl[dd] + s_whist[wave][dd] + loc[k];
#pragma unroll
            for (int w = 0; w < RW; w++) s_rec[pos * RW + w] = r[k][w];
        }
    }
    __syncthreads();

    // write out: consecutive LDS slots of one digit go to consecutive slots
    long long csum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int i = k * BLOCK + tid;
        if (i < count) {
            uint32_t rr[RW];
#pragma unroll
            for (int w = 0; w < RW; w++) rr[w] = s_rec[i * RW + w];
            const uint32_t dd = digit_of<RW>(rr, pd);
            const int64_t o = s_gdelta[dd] + i;
            if (o < sbeg || o >= send) {  // only reachable with inconsistent counts
                atomicOr(error_flag, 4u);
                continue;
            }
            RecIO<RW>::store(dst, o, rr);
            if (checksum_out) {
#pragma unroll
                for (int w = 0; w < RW; w++) csum += sbyte_sum(rr[w]);
            }
        }
    }
    if (checksum_out) {
        csum = wave_sum_ll(csum);
        if (lane == 0) s_sum[wave] = csum;
        __syncthreads();
        if (tid == 0) {
            long long a = 0;
            for (int w = 0; w < NW; w++) a += s_sum[w];
            atomicAdd((unsigned long long*)checksum_out, (unsigned long long)a);
        }
    }
}

// ---------------------------------------------------------------------------
template <int RW>
__global__ __launch_bounds__(kBlock) void copy_kernel(const uint32_t* __restrict__ src,
                                                      uint32_t* __restrict__ dst, int64_t n,
                                                      long long* __restrict__ checksum) {
    __shared__ long long s_sum[kWaves];
    long long csum = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        uint32_t r[RW];
        RecIO<RW>::load(src, i, r);
        RecIO<RW>::store(dst, i, r);
        if (checksum) {
#pragma unroll
            for (int w = 0; w < RW; w++) csum += sbyte_sum(r[w]);
        }
    }
    if (checksum) {
        csum = wave_sum_ll(csum);
        if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = csum;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long a = 0;
            for (int w = 0; w < kWaves; w++) a += s_sum[w];
            atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
        }
    }
}

// copy of the records inside a plan's segments only (one block per tile);
// used when the segments are scattered through the buffer
template <int RW>
__global__ __launch_bounds__(kBlock) void seg_copy_kernel(const uint32_t* __restrict__ src,
                                                          uint32_t* __restrict__ dst, SegDev seg,
                                                          int tile_recs, long long* __restrict__ checksum) {
    __shared__ long long s_sum[kWaves];
    const int64_t tile = blockIdx.x;
    const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
    const int64_t sbeg = seg.begin[s];
    const int64_t begin = sbeg + (tile - seg.tile_start[s]) * tile_recs;
    const int64_t end = min(sbeg + seg.count[s], begin + (int64_t)tile_recs);
    long long csum = 0;
    for (int64_t i = begin + threadIdx.x; i < end; i += kBlock) {
        uint32_t r[RW];
        RecIO<RW>::load(src, i, r);
        RecIO<RW>::store(dst, i, r);
        if (checksum) {
#pragma unroll
            for (int w = 0; w < RW; w++) csum += sbyte_sum(r[w]);
        }
    }
    if (checksum) {
        csum = wave_sum_ll(csum);
        if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = csum;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long a = 0;
            for (int w = 0; w < kWaves; w++) a += s_sum[w];
            atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
        }
    }
}

__global__ void add_u32_kernel(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

__global__ __launch_bounds__(256) void fanout_kernel(FanOut f) {
    const FanEntry e = f.e[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < e.nwords; i += blockDim.x) e.dst[i] = e.src ? e.src[i] : e.fill;
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// host side

// Tile shapes compiled for every record width: 256 and 512 threads with
// hi or hi/2 records per thread, 1024 threads with hi/2 or hi/4, where hi
// fills 64 KB of LDS at 256 threads.
static int items_hi(int rw) { return rw == 1 ? 32 : (rw == 2 ? 32 : (rw == 4 ? 16 : 8)); }

bool tile_supported(int rw, TileCfg c) {
    const int hi = items_hi(rw);
    if (c.block == 256 || c.block == 512) return c.items == hi || c.items == hi / 2;
    if (c.block == 1024) return c.items == hi / 2 || c.items == hi / 4;
    return false;
}

// 1024 threads x hi/2 records (8192-record tiles for 16-byte records) was the
// fastest shape measured on MI355X for 16-byte records (one workgroup per CU,
// 16 waves); MPSORT_GPU_TILE="block:items" overrides it.
TileCfg default_tile(int rw) {
    TileCfg c{1024, items_hi(rw) / 2};
    const char* e = getenv("MPSORT_GPU_TILE");
    if (e) {
        int b = 0, it = 0;
        if (sscanf(e, "%d:%d", &b, &it) == 2 && tile_supported(rw, TileCfg{b, it})) c = TileCfg{b, it};
    }
    return c;
}

int launch_add_u32(uint32_t* p, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(add_u32_kernel, dim3(1), dim3(64), 0, st, p, v);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, p, v, n);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_fanout(const FanOut& f, hipStream_t st) {
    if (f.n <= 0) return 0;
    if (f.n > kFanMax) return -1;
    hipLaunchKernelGGL(fanout_kernel, dim3((unsigned)f.n), dim3(256), 0, st, f);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_copy(const void* src, void* dst, int64_t nrec, int rw, long long* checksum,
                hipStream_t st) {
    if (nrec <= 0) return 0;
    int64_t blocks = (nrec + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    dim3 g((unsigned)blocks), b(kBlock);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    switch (rw) {
        case 1: hipLaunchKernelGGL(copy_kernel<1>, g, b, 0, st, s, d, nrec, checksum); break;
        case 2: hipLaunchKernelGGL(copy_kernel<2>, g, b, 0, st, s, d, nrec, checksum); break;
        case 4: hipLaunchKernelGGL(copy_kernel<4>, g, b, 0, st, s, d, nrec, checksum); break;
        case 8: hipLaunchKernelGGL(copy_kernel<8>, g, b, 0, st, s, d, nrec, checksum); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

// copy src -> dst over the plan's segments: the whole range when they tile
// [0, nrec), tile by tile otherwise
static int plan_copy(const SegSortPlan& pl, const void* src, void* dst, long long* checksum,
                     hipStream_t st) {
    if (!pl.sparse) return launch_copy(src, dst, pl.nrec, pl.rw, checksum, st);
    if (pl.ntiles <= 0) return 0;
    SegDev sd{pl.d_begin, pl.d_count, pl.d_tile_start, pl.nseg, 0};
    dim3 g((unsigned)pl.ntiles), b(kBlock);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    const int tr = pl.tile.block * pl.tile.items;
    switch (pl.rw) {
        case 1: hipLaunchKernelGGL(seg_copy_kernel<1>, g, b, 0, st, s, d, sd, tr, checksum); break;
        case 2: hipLaunchKernelGGL(seg_copy_kernel<2>, g, b, 0, st, s, d, sd, tr, checksum); break;
        case 4: hipLaunchKernelGGL(seg_copy_kernel<4>, g, b, 0, st, s, d, sd, tr, checksum); break;
        case 8: hipLaunchKernelGGL(seg_copy_kernel<8>, g, b, 0, st, s, d, sd, tr, checksum); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

template <int RW>
static void launch_hist(const SegSortPlan& pl, const void* src, int npass, const PassTable& pt,
                        long long* cs, hipStream_t st) {
    SegDev sd{pl.d_begin, pl.d_count, pl.d_tile_start, pl.nseg, 0};
    // about 2048 blocks (8 per CU), each over a contiguous run of tiles
    const int64_t tpb = (pl.ntiles + 2047) / 2048;
    const int64_t blocks = (pl.ntiles + tpb - 1) / tpb;
    hipLaunchKernelGGL(hist_kernel<RW>, dim3((unsigned)blocks), dim3(kBlock), (size_t)npass * kBins * sizeof(uint32_t), st,
                       (const uint32_t*)src, sd, pl.ntiles, tpb, pl.tile.block * pl.tile.items, npass,
                       pt, consecutive_key_bytes(pt, npass), pl.d_hist, cs);
}

template <int RW, int BLOCK, int ITEMS>
static void launch_pass_t(const SegSortPlan& pl, const void* src, void* dst, int npass, int pidx,
                          const PassDesc& pd, uint32_t* ticket, uint32_t epoch, long long* cs,
                          uint32_t* err, hipStream_t st) {
    SegDev sd{pl.d_begin, pl.d_count, pl.d_tile_start, pl.nseg, 0};
    hipLaunchKernelGGL((onesweep_kernel<RW, BLOCK, ITEMS>), dim3((unsigned)pl.ntiles), dim3(BLOCK), 0, st,
                       (const uint32_t*)src, (uint32_t*)dst, sd, pl.d_digit_base, npass, pidx, pd,
                       pl.d_lookback, ticket, epoch, pl.d_epoch_base, cs, err);
}

template <int RW>
static int launch_pass(const SegSortPlan& pl, const void* src, void* dst, int npass, int pidx,
                       const PassDesc& pd, uint32_t* ticket, uint32_t epoch, long long* cs,
                       uint32_t* err, hipStream_t st) {
    constexpr int HI = RW == 1 ? 32 : (RW == 2 ? 32 : (RW == 4 ? 16 : 8));
    constexpr int LO = HI / 2;
    const TileCfg c = pl.tile;
    if (c.block == 256 && c.items == HI) launch_pass_t<RW, 256, HI>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else if (c.block == 256 && c.items == LO) launch_pass_t<RW, 256, LO>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else if (c.block == 512 && c.items == LO) launch_pass_t<RW, 512, LO>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else if (c.block == 512 && c.items == HI) launch_pass_t<RW, 512, HI>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else if (c.block == 1024 && c.items == LO) launch_pass_t<RW, 1024, LO>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else if (c.block == 1024 && c.items == HI / 4) launch_pass_t<RW, 1024, HI / 4>(pl, src, dst, npass, pidx, pd, ticket, epoch, cs, err, st);
    else return -1;
    return 0;
}

int seg_clear(const SegSortPlan& pl, hipStream_t st) {
    const int npass = pl.ks.key_bytes;
    int rc = launch_fill_u32(pl.d_hist, 0u, (int64_t)pl.nseg * npass * kBins, st);
    if (rc) return rc;
    rc = launch_fill_u32(pl.d_pass_flags, 0u, kMaxPasses + 1, st);
    if (rc) return rc;
    return launch_fill_u32(pl.d_tickets, 0u, kMaxPasses, st);
}

int seg_count(const SegSortPlan& pl, const void* src, long long* cs_in, hipStream_t st) {
    const int npass = pl.ks.key_bytes;
    if (npass <= 0 || npass > kMaxPasses) return -1;
    int rc = seg_clear(pl, st);
    if (rc) return rc;
    if (pl.nrec == 0 || pl.ntiles == 0) return 0;
    const PassTable pt = make_pass_table(pl.ks);
    switch (pl.rw) {
        case 1: launch_hist<1>(pl, src, npass, pt, cs_in, st); break;
        case 2: launch_hist<2>(pl, src, npass, pt, cs_in, st); break;
        case 4: launch_hist<4>(pl, src, npass, pt, cs_in, st); break;
        case 8: launch_hist<8>(pl, src, npass, pt, cs_in, st); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int seg_passes(const SegSortPlan& pl, const SortBuffers& bufs, hipStream_t st, long long* cs_out,
               int* passes_run, hipEvent_t* pass_events) {
    *passes_run = 0;
    const int npass = pl.ks.key_bytes;
    uint32_t* err = (uint32_t*)(pl.d_pass_flags + kMaxPasses);
    if (pl.nrec == 0 || pl.ntiles == 0) return 0;
    if (npass <= 0 || npass > kMaxPasses) return -1;
    if (!tile_supported(pl.rw, pl.tile)) return -1;
    const PassTable pt = make_pass_table(pl.ks);
    hipLaunchKernelGGL(scan_kernel, dim3((unsigned)((int64_t)pl.nseg * npass)), dim3(kBlock), 0, st,
                       pl.d_hist, pl.d_begin, npass, pl.d_digit_base, pl.d_pass_flags);
    MPSG_CHECK(hipGetLastError());

    int todo[kMaxPasses], k = 0;
    if (pl.all_passes) {
        // no read-back (the launch sequence may be captured in a graph)
        for (int b = 0; b < npass; b++) todo[k++] = b;
    } else {
        uint32_t flags[kMaxPasses];
        MPSG_CHECK(hipMemcpyAsync(flags, pl.d_pass_flags, sizeof(uint32_t) * npass,
                                  hipMemcpyDeviceToHost, st));
        MPSG_CHECK(hipStreamSynchronize(st));
        for (int b = 0; b < npass; b++)
            if (flags[b]) todo[k++] = b;
    }

    void* src = bufs.src;
    void* X = bufs.dst;
    void* Y = bufs.tmp;
    int rc = 0;
    if (k == 0) {
        if (src != X) return plan_copy(pl, src, X, cs_out, st);
        return 0;
    }
    const void* cur = src;
    if ((k & 1) && src == X) {
        rc = plan_copy(pl, src, Y, nullptr, st);
        if (rc) return rc;
        cur = Y;
    }
    int dev = 0;
    MPSG_CHECK(hipGetDevice(&dev));
    for (int i = 0; i < k; i++) {
        void* out = ((k - 1 - i) % 2 == 0) ? X : Y;
        const uint32_t epoch = pl.d_epoch_base ? (uint32_t)(i + 1) : next_epoch(dev);
        const PassDesc pd = pt.p[todo[i]];
        long long* cso = (i == k - 1) ? cs_out : nullptr;
        if (pass_events) MPSG_CHECK(hipEventRecord(pass_events[2 * i], st));
        switch (pl.rw) {
            case 1: rc = launch_pass<1>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 2: rc = launch_pass<2>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 4: rc = launch_pass<4>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 8: rc = launch_pass<8>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            default: rc = -1;
        }
        if (rc) return rc;
        MPSG_CHECK(hipGetLastError());
        if (pass_events) MPSG_CHECK(hipEventRecord(pass_events[2 * i + 1], st));
        cur = out;
    }
    *passes_run = k;
    return 0;
}

int segmented_sort(const SegSortPlan& pl, const SortBuffers& bufs, hipStream_t st,
                   bool checksum_in, bool checksum_out, int* passes_run, hipEvent_t* pass_events) {
    *passes_run = 0;
    int rc = seg_count(pl, bufs.src, checksum_in ? pl.d_checksum : nullptr, st);
    if (rc) return rc;
    return seg_passes(pl, bufs, st, checksum_out ? pl.d_checksum + 1 : nullptr, passes_run, pass_events);
}

}  // namespace mpsg
