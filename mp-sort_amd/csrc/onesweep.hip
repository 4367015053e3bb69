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
//   * tiles of 256 threads x ITEMS records, tile ids from an atomic ticket so
//     every predecessor tile is already resident (forward progress for the
//     look-back);
//   * wave-striped loads (item k of lane l is record k*64+l of the wave's
//     slice), so ranking items in (k, lane) order is input order: stable;
//   * per-item digit ranking by 8 wave ballots (match-any), per-wave digit
//     counters in LDS;
//   * decoupled look-back over the tiles of the same segment, one thread per
//     digit, 64-bit flag words tagged with a per-launch epoch;
//   * records are staged in LDS in digit order and written out in runs that
//     are contiguous per digit (coalesced).
#include "device.h"

namespace mpsg {

struct PassTable {
    PassDesc p[kMaxPasses];
};

// ---------------------------------------------------------------------------
// upfront histogram of all digit places
template <int RW>
__global__ __launch_bounds__(kBlock) void hist_kernel(const uint32_t* __restrict__ src, SegDev seg,
                                                      int npass, PassTable pt,
                                                      uint32_t* __restrict__ hist,
                                                      long long* __restrict__ checksum) {
    constexpr int ITEMS = TileShape<RW>::kItems;
    constexpr int TILE = kBlock * ITEMS;
    __shared__ uint32_t s_hist[kMaxPasses * kBins];
    __shared__ long long s_sum[kWaves];
    const int tid = threadIdx.x;
    for (int i = tid; i < npass * kBins; i += kBlock) s_hist[i] = 0;
    __syncthreads();

    const int64_t tile = blockIdx.x;
    const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
    const int64_t t = tile - seg.tile_start[s];
    const int64_t sbeg = seg.begin[s];
    const int64_t begin = sbeg + t * TILE;
    const int64_t end = min(sbeg + seg.count[s], begin + (int64_t)TILE);

    long long csum = 0;
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + (int64_t)k * kBlock + tid;
        if (i < end) {
            uint32_t r[RW];
            RecIO<RW>::load(src, i, r);
            if (checksum) {
#pragma unroll
                for (int w = 0; w < RW; w++) csum += sbyte_sum(r[w]);
            }
            for (int p = 0; p < npass; p++) {
                const uint32_t d = digit_of<RW>(r, pt.p[p]);
                atomicAdd(&s_hist[p * kBins + d], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* h = hist + (int64_t)s * npass * kBins;
    for (int i = tid; i < npass * kBins; i += kBlock) {
        const uint32_t c = s_hist[i];
        if (c) atomicAdd(&h[i], c);
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

// ---------------------------------------------------------------------------
// one onesweep digit pass
template <int RW>
__global__ __launch_bounds__(kBlock) void onesweep_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, SegDev seg,
    const int64_t* __restrict__ digit_base, int npass, int pass_idx, PassDesc pd,
    uint64_t* __restrict__ lookback, uint32_t* __restrict__ ticket, uint32_t epoch,
    long long* __restrict__ checksum_out, uint32_t* __restrict__ error_flag) {
    constexpr int ITEMS = TileShape<RW>::kItems;
    constexpr int TILE = kBlock * ITEMS;
    constexpr int WAVE_TILE = 64 * ITEMS;
    __shared__ __attribute__((aligned(16))) uint32_t s_rec[TILE * RW];
    __shared__ uint32_t s_whist[kWaves][kBins];
    __shared__ uint32_t s_excl[kBins];
    __shared__ int64_t s_gdelta[kBins];
    __shared__ int64_t s_tile;
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ long long s_sum[kWaves];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) s_tile = (int64_t)atomicAdd(ticket, 1u);
    for (int i = tid; i < kWaves * kBins; i += kBlock) (&s_whist[0][0])[i] = 0;
    __syncthreads();

    const int64_t tile = s_tile;
    const int s = upper_index<int64_t>(seg.tile_start, seg.nseg, tile);
    const int64_t t = tile - seg.tile_start[s];
    const int64_t sbeg = seg.begin[s];
    const int64_t begin = sbeg + t * TILE;
    const int64_t end = min(sbeg + seg.count[s], begin + (int64_t)TILE);
    const int count = (int)(end - begin);

    uint32_t r[ITEMS][RW];
    uint32_t dg[ITEMS];
    uint32_t loc[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = begin + wave * WAVE_TILE + k * 64 + lane;
        if (i < end) {
            RecIO<RW>::load(src, i, r[k]);
            dg[k] = digit_of<RW>(r[k], pd);
        } else {
#pragma unroll
            for (int w = 0; w < RW; w++) r[k][w] = 0;
            dg[k] = kBins;  // not a record
        }
    }

    // rank inside the wave: items in (k, lane) order are in input order
    volatile uint32_t* wh = s_whist[wave];
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
        if (valid && rank == 0) wh[d] = base + cnt;
    }
    __syncthreads();

    // thread d owns digit d: wave offsets, tile-local digit start, look-back
    const int d = tid;
    uint32_t total = 0;
    {
        uint32_t c[kWaves];
#pragma unroll
        for (int w = 0; w < kWaves; w++) c[w] = s_whist[w][d];
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) { s_whist[w][d] = acc; acc += c[w]; }
        total = acc;
    }
    const uint32_t excl = block_exclusive_scan(total, s_tmp, nullptr);
    s_excl[d] = excl;

    int64_t prefix = 0;
    uint64_t* my = lookback + tile * kBins + d;
    if (t == 0) {
        store_agent(my, make_flag(epoch, 1, total));
    } else {
        store_agent(my, make_flag(epoch, 0, total));
        int64_t j = tile - 1;
        uint32_t spins = 0;
        while (true) {
            const uint64_t f = load_agent(lookback + j * kBins + d);
            if ((uint32_t)(f >> 41) != epoch) {
                if (++spins > kSpinLimit) {
                    atomicOr(error_flag, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += (int64_t)(f & kFlagValueMask);
            if ((f >> 40) & 1ull) break;
            --j;
        }
        store_agent(my, make_flag(epoch, 1, (uint64_t)(prefix + total)));
    }
    s_gdelta[d] = digit_base[((int64_t)s * npass + pass_idx) * kBins + d] + prefix - (int64_t)excl;
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
    for (int i = tid; i < count; i += kBlock) {
        uint32_t rr[RW];
#pragma unroll
        for (int w = 0; w < RW; w++) rr[w] = s_rec[i * RW + w];
        const uint32_t dd = digit_of<RW>(rr, pd);
        const int64_t o = s_gdelta[dd] + i;
        if (o < sbeg || o >= sbeg + seg.count[s]) {  // only reachable with inconsistent counts
            atomicOr(error_flag, 4u);
            continue;
        }
        RecIO<RW>::store(dst, o, rr);
        if (checksum_out) {
#pragma unroll
            for (int w = 0; w < RW; w++) csum += sbyte_sum(rr[w]);
        }
    }
    if (checksum_out) {
        csum = wave_sum_ll(csum);
        if (lane == 0) s_sum[wave] = csum;
        __syncthreads();
        if (tid == 0) {
            long long a = 0;
            for (int w = 0; w < kWaves; w++) a += s_sum[w];
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

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// host side

int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, p, v, n);
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

template <int RW>
static void launch_hist(const SegSortPlan& pl, const void* src, int npass, const PassTable& pt,
                        long long* cs, hipStream_t st) {
    SegDev sd{pl.d_begin, pl.d_count, pl.d_tile_start, pl.nseg, 0};
    hipLaunchKernelGGL(hist_kernel<RW>, dim3((unsigned)pl.ntiles), dim3(kBlock), 0, st,
                       (const uint32_t*)src, sd, npass, pt, pl.d_hist, cs);
}

template <int RW>
static void launch_pass(const SegSortPlan& pl, const void* src, void* dst, int npass, int pidx,
                        const PassDesc& pd, uint32_t* ticket, uint32_t epoch, long long* cs,
                        uint32_t* err, hipStream_t st) {
    SegDev sd{pl.d_begin, pl.d_count, pl.d_tile_start, pl.nseg, 0};
    hipLaunchKernelGGL(onesweep_kernel<RW>, dim3((unsigned)pl.ntiles), dim3(kBlock), 0, st,
                       (const uint32_t*)src, (uint32_t*)dst, sd, pl.d_digit_base, npass, pidx, pd,
                       pl.d_lookback, ticket, epoch, cs, err);
}

int segmented_sort(const SegSortPlan& pl, const SortBuffers& bufs, hipStream_t st,
                   bool checksum_in, bool checksum_out, int* passes_run, hipEvent_t* pass_events) {
    *passes_run = 0;
    const int npass = pl.ks.key_bytes;
    long long* cs_in = checksum_in ? pl.d_checksum : nullptr;
    long long* cs_out = checksum_out ? pl.d_checksum + 1 : nullptr;
    uint32_t* err = (uint32_t*)(pl.d_pass_flags + kMaxPasses);
    if (pl.nrec == 0 || pl.ntiles == 0) return 0;
    if (npass <= 0 || npass > kMaxPasses) return -1;

    PassTable pt;
    for (int b = 0; b < npass; b++) pt.p[b] = make_pass(pl.ks, b);

    // clear counters: hist, pass flags (+ error flag), tickets
    int rc = launch_fill_u32(pl.d_hist, 0u, (int64_t)pl.nseg * npass * kBins, st);
    if (rc) return rc;
    rc = launch_fill_u32(pl.d_pass_flags, 0u, kMaxPasses + 1, st);
    if (rc) return rc;
    rc = launch_fill_u32(pl.d_tickets, 0u, kMaxPasses, st);
    if (rc) return rc;

    switch (pl.rw) {
        case 1: launch_hist<1>(pl, bufs.src, npass, pt, cs_in, st); break;
        case 2: launch_hist<2>(pl, bufs.src, npass, pt, cs_in, st); break;
        case 4: launch_hist<4>(pl, bufs.src, npass, pt, cs_in, st); break;
        case 8: launch_hist<8>(pl, bufs.src, npass, pt, cs_in, st); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(scan_kernel, dim3((unsigned)((int64_t)pl.nseg * npass)), dim3(kBlock), 0, st,
                       pl.d_hist, pl.d_begin, npass, pl.d_digit_base, pl.d_pass_flags);
    MPSG_CHECK(hipGetLastError());

    uint32_t flags[kMaxPasses];
    MPSG_CHECK(hipMemcpyAsync(flags, pl.d_pass_flags, sizeof(uint32_t) * npass,
                              hipMemcpyDeviceToHost, st));
    MPSG_CHECK(hipStreamSynchronize(st));
    int todo[kMaxPasses], k = 0;
    for (int b = 0; b < npass; b++)
        if (flags[b]) todo[k++] = b;

    void* src = bufs.src;
    void* X = bufs.dst;
    void* Y = bufs.tmp;
    if (k == 0) {
        if (src != X) return launch_copy(src, X, pl.nrec, pl.rw, cs_out, st);
        return 0;
    }
    const void* cur = src;
    if ((k & 1) && src == X) {
        rc = launch_copy(src, Y, pl.nrec, pl.rw, nullptr, st);
        if (rc) return rc;
        cur = Y;
    }
    int dev = 0;
    MPSG_CHECK(hipGetDevice(&dev));
    for (int i = 0; i < k; i++) {
        void* out = ((k - 1 - i) % 2 == 0) ? X : Y;
        const uint32_t epoch = next_epoch(dev);
        const PassDesc pd = pt.p[todo[i]];
        long long* cso = (i == k - 1) ? cs_out : nullptr;
        if (pass_events) MPSG_CHECK(hipEventRecord(pass_events[2 * i], st));
        switch (pl.rw) {
            case 1: launch_pass<1>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 2: launch_pass<2>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 4: launch_pass<4>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
            case 8: launch_pass<8>(pl, cur, out, npass, todo[i], pd, pl.d_tickets + i, epoch, cso, err, st); break;
        }
        MPSG_CHECK(hipGetLastError());
        if (pass_events) MPSG_CHECK(hipEventRecord(pass_events[2 * i + 1], st));
        cur = out;
    }
    *passes_run = k;
    return 0;
}

}  // namespace mpsg
