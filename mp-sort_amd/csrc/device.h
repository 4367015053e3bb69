// device.h -- small device helpers shared by the kernels (gfx950, wave64).
#pragma once
#include "internal.h"

namespace mpsg {

// Look-back flag word: [epoch:23][kind:1][value:40].  A word whose epoch is
// not the current launch's epoch is "not ready"; kind 1 = inclusive prefix.
constexpr int kFlagValueBits = 40;
constexpr uint64_t kFlagValueMask = (1ull << kFlagValueBits) - 1;
constexpr uint32_t kEpochLimit = 1u << 23;
// a look-back wait gives up after this many polls and raises an error flag
constexpr uint32_t kSpinLimit = 1u << 24;
// predecessor flags read per look-back step
constexpr int kLookback = 4;

__device__ __forceinline__ uint64_t make_flag(uint32_t epoch, uint32_t inclusive, uint64_t v) {
    return ((uint64_t)epoch << 41) | ((uint64_t)inclusive << 40) | (v & kFlagValueMask);
}

__device__ __forceinline__ void store_agent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t load_agent(uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// largest s in [0, n) with starts[s] <= x (starts non-decreasing, starts[0] <= x)
template <typename T>
__device__ __forceinline__ int upper_index(const T* starts, int n, T x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (starts[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// sum of the 4 bytes of w read as signed chars (mpsort-mpi.c:149-159 checksum)
__device__ __forceinline__ int sbyte_sum(uint32_t w) {
    uint32_t u = __builtin_amdgcn_sad_u8(w, 0u, 0u);
    return (int)u - 256 * (int)__popc(w & 0x80808080u);
}

// r[widx] for a wave-uniform widx, written as mask arithmetic: a select
// between array elements can be folded into a dynamically indexed load, which
// would move the whole record array to scratch memory
template <int RW>
__device__ __forceinline__ uint32_t pick_word(const uint32_t (&r)[RW], int widx) {
    uint32_t w = 0;
#pragma unroll
    for (int j = 0; j < RW; j++) w |= r[j] & (0u - (uint32_t)(widx == j));
    return w;
}

template <int RW>
__device__ __forceinline__ uint32_t digit_of(const uint32_t (&r)[RW], const PassDesc& pd) {
    return ((pick_word<RW>(r, pd.widx) >> pd.shift) & 0xffu) ^ pd.xmask;
}

// Counts the digit of every pass of record r in h[p * kBins + digit] (LDS
// or global counters).  b0 >= 0: the passes are the record's bytes b0,
// b0 + 1, ... (at most 8 of them, consecutive_key_bytes), read as two words
// once; otherwise each pass picks its own word.
template <int RW>
__device__ __forceinline__ void count_digits(const uint32_t (&r)[RW], int npass, const PassTable& pt, int b0,
                                             uint32_t* h) {
    if (b0 >= 0) {
        const int w = b0 >> 2, sh = (b0 & 3) * 8;
        const uint32_t a = pick_word<RW>(r, w), b = pick_word<RW>(r, w + 1), c = pick_word<RW>(r, w + 2);
        const uint32_t lo = sh ? ((a >> sh) | (b << (32 - sh))) : a;  // bytes b0 .. b0 + 3
        const uint32_t hi = sh ? ((b >> sh) | (c << (32 - sh))) : b;  // bytes b0 + 4 .. b0 + 7
#pragma unroll
        for (int p = 0; p < 8; p++) {
            if (p < npass) {
                const uint32_t d = (((p < 4 ? lo : hi) >> (8 * (p & 3))) & 0xffu) ^ pt.p[p].xmask;
                atomicAdd(&h[p * kBins + d], 1u);
            }
        }
    } else {
        for (int p = 0; p < npass; p++) atomicAdd(&h[p * kBins + digit_of<RW>(r, pt.p[p])], 1u);
    }
}

template <int RW> struct RecIO;
template <> struct RecIO<1> {
    __device__ static void load(const uint32_t* p, int64_t i, uint32_t (&r)[1]) { r[0] = p[i]; }
    __device__ static void store(uint32_t* p, int64_t i, const uint32_t (&r)[1]) { p[i] = r[0]; }
};
template <> struct RecIO<2> {
    __device__ static void load(const uint32_t* p, int64_t i, uint32_t (&r)[2]) {
        uint2 v = reinterpret_cast<const uint2*>(p)[i];
        r[0] = v.x; r[1] = v.y;
    }
    __device__ static void store(uint32_t* p, int64_t i, const uint32_t (&r)[2]) {
        reinterpret_cast<uint2*>(p)[i] = make_uint2(r[0], r[1]);
    }
};
template <> struct RecIO<4> {
    __device__ static void load(const uint32_t* p, int64_t i, uint32_t (&r)[4]) {
        uint4 v = reinterpret_cast<const uint4*>(p)[i];
        r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
    }
    __device__ static void store(uint32_t* p, int64_t i, const uint32_t (&r)[4]) {
        reinterpret_cast<uint4*>(p)[i] = make_uint4(r[0], r[1], r[2], r[3]);
    }
};
template <> struct RecIO<8> {
    __device__ static void load(const uint32_t* p, int64_t i, uint32_t (&r)[8]) {
        const uint4* q = reinterpret_cast<const uint4*>(p) + 2 * i;
        uint4 a = q[0], b = q[1];
        r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
        r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
    }
    __device__ static void store(uint32_t* p, int64_t i, const uint32_t (&r)[8]) {
        uint4* q = reinterpret_cast<uint4*>(p) + 2 * i;
        q[0] = make_uint4(r[0], r[1], r[2], r[3]);
        q[1] = make_uint4(r[4], r[5], r[6], r[7]);
    }
};

// wave-wide sum of a 64-bit value
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block (256 threads) exclusive scan; s_tmp needs kWaves entries.  Also
// returns the block total.  Contains two __syncthreads.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        uint32_t t = s_tmp[w];
        pre += (w < wave) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + x - v;
}

}  // namespace mpsg
