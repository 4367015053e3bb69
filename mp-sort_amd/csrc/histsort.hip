// histsort.hip -- the splitter search, layout solve and exchange of the
// histogram sort (mpsort-mpi.c:333-604), with the P MPI ranks of the
// reference held as P virtual-rank segments of one GPU's memory.
//
//   bisect_kernel   _find_Pmax_Pmin_C (mpsort-mpi.c:606-661) + the piter
//                   bisection loop (internal-parallel.h:128-248,
//                   mpsort-mpi.c:381-441) + _histogram counts
//                   (internal-parallel.h:97-126).  One wave per splitter, one
//                   lane per segment; every round does two binary searches
//                   per segment inside an index window that only shrinks.
//   layout_kernel   _solve_for_layout_mpi (mpsort-mpi.c:663-727), one thread
//                   per receiving rank.
//   runs_kernel     SendCount / RecvDispl (mpsort-mpi.c:483-510) as a table of
//                   (sender, receiver) runs in output order, and the count of
//                   non-empty runs per receiver and per sender.
//   runs_vote_kernel, runs_compact_kernel  the dense / sparse choice of
//                   MPIU_Alltoallv (mp-mpiu.c:125-136) and the table of the
//                   non-empty runs alone (the sparse form, mp-mpiu.c:158-236,
//                   which skips the empty pairs)
//   exchange_kernel MPIU_Alltoallv (mp-mpiu.c:69-156): every output record is
//                   copied from its sender run, found by a search in the
//                   chosen run table; receiver buffers hold the sender runs
//                   in sender order.
#include "device.h"
#include "model_dev.h"

namespace mpsg {

struct Key {
    uint64_t w[4];  // little-endian: w[0] least significant
};

__device__ __forceinline__ Key load_key(const uint32_t* __restrict__ recs, int rw, int64_t idx,
                                        const KeySpec& ks) {
    Key k = {{0, 0, 0, 0}};
    const uint32_t* rec = recs + idx * rw;
    if (ks.key_bytes == 8 && (ks.key_off & 3) == 0) {
        const uint32_t lo = rec[ks.key_off >> 2] ^ ((uint32_t)ks.xmask[0] | ((uint32_t)ks.xmask[1] << 8) |
                                                    ((uint32_t)ks.xmask[2] << 16) | ((uint32_t)ks.xmask[3] << 24));
        const uint32_t hi = rec[(ks.key_off >> 2) + 1] ^ ((uint32_t)ks.xmask[4] | ((uint32_t)ks.xmask[5] << 8) |
                                                          ((uint32_t)ks.xmask[6] << 16) | ((uint32_t)ks.xmask[7] << 24));
        k.w[0] = (uint64_t)lo | ((uint64_t)hi << 32);
        return k;
    }
    if (ks.key_bytes == 4 && (ks.key_off & 3) == 0) {
        const uint32_t v = rec[ks.key_off >> 2] ^ ((uint32_t)ks.xmask[0] | ((uint32_t)ks.xmask[1] << 8) |
                                                   ((uint32_t)ks.xmask[2] << 16) | ((uint32_t)ks.xmask[3] << 24));
        k.w[0] = v;
        return k;
    }
    const uint8_t* p = reinterpret_cast<const uint8_t*>(rec) + ks.key_off;
    for (int b = 0; b < ks.key_bytes; b++) {
        const uint64_t byte = (uint64_t)(uint8_t)(p[b] ^ ks.xmask[b]);
        k.w[b >> 3] |= byte << ((b & 7) * 8);
    }
    return k;
}

__device__ __forceinline__ int key_cmp(const Key& a, const Key& b, int nw) {
#pragma unroll
    for (int w = 3; w >= 0; w--) {
        if (w < nw) {
            if (a.w[w] < b.w[w]) return -1;
            if (a.w[w] > b.w[w]) return 1;
        }
    }
    return 0;
}

// floor((a+b)/2) as (nw*64+1)-bit arithmetic: radixsort.c:55-60, 111-144
__device__ __forceinline__ Key key_mid(const Key& a, const Key& b, int nw) {
    Key s;
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint64_t t = a.w[w] + carry;
        const uint64_t c1 = t < carry;
        const uint64_t u = t + b.w[w];
        const uint64_t c2 = u < t;
        s.w[w] = (w < nw) ? u : 0;
        carry = (w < nw) ? (c1 | c2) : carry;
    }
    Key m;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint64_t hi = (w + 1 < nw) ? s.w[w + 1] : ((w + 1 == nw) ? carry : 0);
        m.w[w] = (w < nw) ? ((s.w[w] >> 1) | (hi << 63)) : 0;
    }
    return m;
}

__device__ __forceinline__ Key shfl_key(const Key& k, int off) {
    Key o;
#pragma unroll
    for (int w = 0; w < 4; w++) o.w[w] = __shfl_xor(k.w[w], off, 64);
    return o;
}

__device__ __forceinline__ Key bcast_key(const Key& k, int src) {
    Key o;
#pragma unroll
    for (int w = 0; w < 4; w++) o.w[w] = __shfl(k.w[w], src, 64);
    return o;
}

// total records over the segments (every lane gets it)
__device__ __forceinline__ int64_t scnt_total(const int64_t* __restrict__ scnt, int P) {
    long long t = 0;
    for (int j = threadIdx.x & 63; j < P; j += 64) t += scnt[j];
    return (int64_t)wave_sum_ll(t);
}

// ---------------------------------------------------------------------------
// Window slots per (splitter, segment): [lo, hi) of the segment's candidate
// records, then #{< v} and #{<= v} of the value v tested last (two values in
// the learned search: slots 2-3 and 4-5).
constexpr int kWinSlots = 6;

// #{< v} and #{<= v} in records [a, b) of a sorted segment; the two searches
// advance together, so their loads overlap
__device__ __forceinline__ void count_in(const uint32_t* __restrict__ recs, int rw, const KeySpec& ks, int64_t base,
                                         int64_t a, int64_t b, const Key& v, int nw, int64_t* lt, int64_t* le) {
    int64_t a1 = a, b1 = b, a2 = a, b2 = b;
    while (a1 < b1 || a2 < b2) {
        const int64_t m1 = a1 + ((b1 - a1) >> 1), m2 = a2 + ((b2 - a2) >> 1);
        const bool s1 = a1 < b1, s2 = a2 < b2;
        const Key k1 = load_key(recs, rw, base + (s1 ? m1 : m2), ks);
        const Key k2 = load_key(recs, rw, base + (s2 ? m2 : m1), ks);
        if (s1) {
            if (key_cmp(k1, v, nw) < 0) a1 = m1 + 1; else b1 = m1;
        }
        if (s2) {
            if (key_cmp(k2, v, nw) <= 0) a2 = m2 + 1; else b2 = m2;
        }
    }
    *lt = a1;
    *le = a2;
}

// The value of rank r (0-based) among the candidates of the windows, when
// there are at most 64 of them (wsum; `mine` this lane's share): gathered
// into LDS, ranked by the wave.  Returns false when r is out of range.
__device__ __forceinline__ bool select_direct(const uint32_t* __restrict__ recs, int rw, const KeySpec& ks,
                                              const int64_t* __restrict__ sbeg, int P, const int64_t* wi,
                                              long long wsum, long long mine, long long r, int nw, Key* s_cand,
                                              Key* out) {
    const int lane = threadIdx.x & 63;
    long long off = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(off, o, 64);
        if (lane >= o) off += y;
    }
    off -= mine;
    for (int j = lane; j < P; j += 64)
        for (int64_t q = wi[kWinSlots * j + 0]; q < wi[kWinSlots * j + 1]; q++)
            s_cand[off++] = load_key(recs, rw, sbeg[j] + q, ks);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bool hit = false;
    Key c = {{0, 0, 0, 0}};
    if (lane < wsum) {
        c = s_cand[lane];
        long long lt = 0, le = 0;
        for (int t = 0; t < (int)wsum; t++) {
            const int cmp = key_cmp(s_cand[t], c, nw);
            lt += cmp < 0 ? 1 : 0;
            le += cmp <= 0 ? 1 : 0;
        }
        hit = lt <= r && r < le;
    }
    const uint64_t hits = __ballot(hit);
    if (!hits) return false;
    *out = bcast_key(c, (int)__ffsll((unsigned long long)hits) - 1);
    return true;
}

__global__ __launch_bounds__(kBlock) void bisect_kernel(
    const uint32_t* __restrict__ recs, int rw, KeySpec ks, const int64_t* __restrict__ sbeg,
    const int64_t* __restrict__ scnt, int P, const int64_t* __restrict__ C,
    int64_t* __restrict__ win, int64_t* __restrict__ myclt, int64_t* __restrict__ mycle,
    uint64_t* __restrict__ split_out, uint64_t* __restrict__ pminmax, int* __restrict__ rounds_out,
    uint32_t* __restrict__ error_flag, ModelDev md, int use_model, int* __restrict__ exact_out) {
    __shared__ Key s_cand[kWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kWaves + wave;  // splitter index
    if (i >= P - 1) return;                   // uniform per wave
    const int nw = (ks.key_bytes + 7) >> 3;
    const int P1 = P + 1;

    // Pmin / Pmax over the non-empty segments (mpsort-mpi.c:642-660)
    Key mn = {{~0ull, ~0ull, ~0ull, ~0ull}}, mx = {{0, 0, 0, 0}};
    bool any = false;
    for (int j = lane; j < P; j += 64) {
        const int64_t n = scnt[j];
        if (n > 0) {
            const Key f = load_key(recs, rw, sbeg[j], ks);
            const Key l = load_key(recs, rw, sbeg[j] + n - 1, ks);
            if (key_cmp(f, mn, nw) < 0) mn = f;
            if (key_cmp(l, mx, nw) > 0) mx = l;
            any = true;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const Key a = shfl_key(mn, off), b = shfl_key(mx, off);
        if (key_cmp(a, mn, nw) < 0) mn = a;
        if (key_cmp(b, mx, nw) > 0) mx = b;
    }
    if (__ballot(any) == 0) {
        mn = Key{{0, 0, 0, 0}};
        mx = mn;
    }
    if (i == 0 && lane == 0) {
        for (int w = 0; w < 4; w++) { pminmax[w] = mn.w[w]; pminmax[4 + w] = mx.w[w]; }
    }

    Key L = mn, R = mx, Pv = mn;
    int rounds = 0, exact = 0;
    const int64_t Ci = C[i + 1];
    int64_t* wi = win + (int64_t)i * P * kWinSlots;
    for (int j = lane; j < P; j += 64) {
        wi[kWinSlots * j + 0] = 0;
        wi[kWinSlots * j + 1] = scnt[j];
    }

    if (use_model && nw == 1 && Ci > 0) {
        // Learned search (a rank of at least 1: C = 0 takes the reference's
        // bisection, which ends next to Pmin).  Every value tested is checked with exact counts,
        // so the result is the reference's splitter k[C-1]; only the values
        // tested differ from bisection.  Round 1 tests the model's proposal
        // F^-1(C / n).  Later rounds test two values together (their four
        // searches per segment overlap): the window-weighted median of every
        // segment's candidate at the target's fractional position f in the
        // windows (a key taken from the data, good for any distribution,
        // duplicates included), and the key interpolated between the
        // bracket's ends at f (good for keys spread evenly, such as a nearly
        // sorted ramp).  When one end has moved twice in a row the target is
        // taken an eighth of the window past its estimate, towards the other
        // end.  An end found too large or too small leaves the window
        // together with every record equal to it.  Once every candidate has
        // one value, or at most 64 remain, the splitter is read off directly.
        bool Lt = false, Rt = false;  // the ends have been tested (and left the windows)
        int moved = 0, repeats = 0;
        bool done = false;
        while (!done) {
            rounds++;
            long long wsum = 0, below = 0;
            Key fk = {{~0ull, 0, 0, 0}}, lk = {{0, 0, 0, 0}};
            for (int j = lane; j < P; j += 64) {
                const int64_t lo = wi[kWinSlots * j + 0], hi = wi[kWinSlots * j + 1];
                wsum += hi - lo;
                below += lo;
                if (hi > lo) {
                    const Key a = load_key(recs, rw, sbeg[j] + lo, ks), b = load_key(recs, rw, sbeg[j] + hi - 1, ks);
                    if (key_cmp(a, fk, nw) < 0) fk = a;
                    if (key_cmp(b, lk, nw) > 0) lk = b;
                }
            }
            const long long mine = wsum;
            wsum = wave_sum_ll(wsum);
            below = wave_sum_ll(below);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const Key a = shfl_key(fk, off), b = shfl_key(lk, off);
                if (key_cmp(a, fk, nw) < 0) fk = a;
                if (key_cmp(b, lk, nw) > 0) lk = b;
            }
            Key v[2];
            int nv = 0;
            if (wsum > 0 && key_cmp(fk, lk, nw) == 0) {
                v[nv++] = fk;  // every candidate has this value
            } else if (wsum > 0 && wsum <= 64) {
                Key c;
                if (select_direct(recs, rw, ks, sbeg, P, wi, wsum, mine, (long long)Ci - 1 - below, nw, s_cand[wave], &c))
                    v[nv++] = c;
            } else if (rounds == 1) {
                double q = ((double)Ci - 0.5) / (double)scnt_total(scnt, P);
                if (use_model == 2) q = q + 0.37 < 1.0 ? q + 0.37 : q - 0.63;  // the fault hook: a wrong proposal
                v[nv++] = Key{{model_inverse_leaves(md, (float)q), 0, 0, 0}};
            } else if (wsum > 0) {
                double t = (double)Ci - 0.5;
                if (repeats >= 2) t += (moved < 0 ? 1.0 : -1.0) * (double)wsum * 0.125;
                double f = (t - (double)below) / (double)wsum;
                f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
                // this lane's candidate (its widest window) and weight
                Key c = {{~0ull, 0, 0, 0}};
                long long wt = 0, best = 0;
                for (int j = lane; j < P; j += 64) {
                    const int64_t lo = wi[kWinSlots * j + 0], hi = wi[kWinSlots * j + 1];
                    wt += hi - lo;
                    if (hi - lo > best) {
                        best = hi - lo;
                        int64_t q = lo + (int64_t)(f * (double)(hi - lo));
                        q = q >= hi ? hi - 1 : q;
                        c = load_key(recs, rw, sbeg[j] + q, ks);
                    }
                }
                // weighted median over the lanes: bitonic sort of (key, weight)
#pragma unroll
                for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
                    for (int jj = k >> 1; jj > 0; jj >>= 1) {
                        const Key oc = shfl_key(c, jj);
                        const long long ow = __shfl_xor(wt, jj, 64);
                        const bool up = (lane & k) == 0, lower = (lane & jj) == 0;
                        const bool take_min = (up == lower);
                        const bool other_less = key_cmp(oc, c, nw) < 0;
                        if (take_min ? other_less : (key_cmp(oc, c, nw) > 0)) {
                            c = oc;
                            wt = ow;
                        }
                    }
                }
                long long cw = wt;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const long long y = __shfl_up(cw, o, 64);
                    if (lane >= o) cw += y;
                }
                const uint64_t med = __ballot(2 * cw >= wsum);
                v[nv++] = bcast_key(c, (int)__ffsll((unsigned long long)med) - 1);
                // the interpolated key
                const double span = (double)(R.w[0] - L.w[0]);
                uint64_t d = (uint64_t)(span * f);
                d = d > R.w[0] - L.w[0] ? R.w[0] - L.w[0] : d;
                v[nv++] = Key{{L.w[0] + d, 0, 0, 0}};
            }
            // keep values that can still be the splitter: inside (L, R), or
            // an end not tested yet
            int kept = 0;
            for (int q = 0; q < nv; q++) {
                const int cl = key_cmp(v[q], L, nw), cr = key_cmp(v[q], R, nw);
                if ((cl > 0 || (cl == 0 && !Lt)) && (cr < 0 || (cr == 0 && !Rt)) &&
                    !(kept == 1 && key_cmp(v[q], v[0], nw) == 0))
                    v[kept++] = v[q];
            }
            nv = kept;
            if (nv == 0) v[nv++] = !Lt ? L : (!Rt ? R : key_mid(L, R, nw));
            // counts of the values in every segment's window
            long long slt[2] = {0, 0}, sle[2] = {0, 0};
            for (int j = lane; j < P; j += 64) {
                int64_t* wj = wi + kWinSlots * j;
                const int64_t base = sbeg[j];
                for (int q = 0; q < nv; q++) {
                    int64_t lt, le;
                    count_in(recs, rw, ks, base, wj[0], wj[1], v[q], nw, &lt, &le);
                    wj[2 + 2 * q] = lt;
                    wj[3 + 2 * q] = le;
                    slt[q] += lt;
                    sle[q] += le;
                }
            }
            int side = 0;
            for (int q = 0; q < nv && !done; q++) {
                slt[q] = wave_sum_ll(slt[q]);
                sle[q] = wave_sum_ll(sle[q]);
                if (slt[q] < Ci && Ci <= sle[q]) {
                    Pv = v[q];
                    done = true;
                    if (rounds == 1) exact = 1;  // the proposal was the splitter
                    if (q == 1)
                        for (int j = lane; j < P; j += 64) {
                            wi[kWinSlots * j + 2] = wi[kWinSlots * j + 4];
                            wi[kWinSlots * j + 3] = wi[kWinSlots * j + 5];
                        }
                }
            }
            if (!done) {
                for (int q = 0; q < nv; q++) {
                    if (slt[q] >= Ci) {
                        // too large: the splitter is below v
                        if (key_cmp(v[q], R, nw) < 0 || !Rt) {
                            R = v[q];
                            Rt = true;
                            side = 1;
                            for (int j = lane; j < P; j += 64) wi[kWinSlots * j + 1] = wi[kWinSlots * j + 2 + 2 * q];
                        }
                    } else {
                        // too small: the splitter is above v
                        if (key_cmp(v[q], L, nw) > 0 || !Lt) {
                            L = v[q];
                            Lt = true;
                            side = -1;
                            for (int j = lane; j < P; j += 64) wi[kWinSlots * j + 0] = wi[kWinSlots * j + 3 + 2 * q];
                        }
                    }
                }
                repeats = (side != 0 && side == moved) ? repeats + 1 : 1;
                moved = side;
            }
            if (rounds > 8 * kMaxKeyBytes + 16) {  // cannot happen for a consistent input
                if (lane == 0) atomicOr(error_flag, 2u);
                break;
            }
        }
    } else {
        // piter_init / piter_bisect / piter_accept (internal-parallel.h:128-248)
        bool narrow = false, stable = false;
        while (!stable) {
            rounds++;
            bool use_found = false;
            Key found = mn;
            if (!narrow) {
                // at most 64 keys left in the windows (the keys in [L, R]):
                // the splitter is the one of rank Ci - 1 - #{< L} among them,
                // found directly; the round below then counts it like any
                // other value
                long long wsum = 0, below = 0;
                for (int j = lane; j < P; j += 64) {
                    wsum += wi[kWinSlots * j + 1] - wi[kWinSlots * j + 0];
                    below += wi[kWinSlots * j + 0];
                }
                const long long mine = wsum;
                wsum = wave_sum_ll(wsum);
                below = wave_sum_ll(below);
                if (wsum > 0 && wsum <= 64)
                    use_found = select_direct(recs, rw, ks, sbeg, P, wi, wsum, mine, (long long)Ci - 1 - below, nw,
                                              s_cand[wave], &found);
            }
            if (use_found) {
                Pv = found;
            } else if (narrow) {
                Pv = R;
                stable = true;
            } else {
                Pv = key_mid(L, R, nw);
                if (key_cmp(Pv, L, nw) <= 0) narrow = true;
            }
            // counts of < Pv and <= Pv per segment, inside the window
            long long slt = 0, sle = 0;
            for (int j = lane; j < P; j += 64) {
                int64_t* wj = wi + kWinSlots * j;
                int64_t lt, le;
                count_in(recs, rw, ks, sbeg[j], wj[0], wj[1], Pv, nw, &lt, &le);
                wj[2] = lt;
                wj[3] = le;
                slt += lt;
                sle += le;
            }
            slt = wave_sum_ll(slt);
            sle = wave_sum_ll(sle);
            if (slt < Ci && Ci <= sle) {
                stable = true;
            } else if (!stable) {
                const bool too_big = slt >= Ci;
                if (too_big) R = Pv; else L = Pv;
                for (int j = lane; j < P; j += 64) {
                    int64_t* wj = wi + kWinSlots * j;
                    if (too_big) wj[1] = wj[3]; else wj[0] = wj[2];
                }
            }
            if (rounds > 8 * kMaxKeyBytes + 16) {  // cannot happen for a consistent input
                if (lane == 0) atomicOr(error_flag, 2u);
                break;
            }
        }
    }
    for (int j = lane; j < P; j += 64) {
        myclt[(int64_t)j * P1 + i + 1] = wi[kWinSlots * j + 2];
        mycle[(int64_t)j * P1 + i + 1] = wi[kWinSlots * j + 3];
    }
    if (lane == 0) {
        for (int w = 0; w < 4; w++) split_out[(int64_t)i * 4 + w] = Pv.w[w];
        rounds_out[i] = rounds;
        if (exact_out) exact_out[i] = exact;
    }
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ long long wave_incl_scan_ll(long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// One wave per receiving rank r (mpsort-mpi.c:663-727), lane j for senders
// j, j + 64, ...: the serial fill from sender 0 takes from sender j
// min(supply_j, max(0, deficit - sum of the earlier supplies)) (a negative
// supply takes nothing and, while some deficit remains, raises the error
// flag, as in the serial loop).  Wave 0 also fills the fixed columns 0 and P
// of the count matrices.
__global__ void layout_kernel(int P, const int64_t* __restrict__ C, const int64_t* __restrict__ scnt,
                              int64_t* __restrict__ myclt, int64_t* __restrict__ mycle,
                              int64_t* __restrict__ myC, int* __restrict__ snz, uint32_t* __restrict__ error_flag) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= P) return;  // uniform per wave
    if (lane == 0) snz[r] = 0;  // counted by runs_kernel
    const int P1 = P + 1;
    const bool last = (r == P - 1);
    long long sure = 0;
    for (int j = lane; j < P; j += 64) sure += last ? scnt[j] : myclt[(int64_t)j * P1 + r + 1];
    sure = wave_sum_ll(sure);
    const long long deficit = C[r + 1] - sure;
    if (deficit < 0 && lane == 0) atomicOr(error_flag, 4u);
    long long before = 0;  // supplies of the senders before this chunk
    for (int j0 = 0; j0 < P; j0 += 64) {
        const int j = j0 + lane;
        int64_t lt = 0, le = 0;
        if (j < P) {
            lt = last ? scnt[j] : myclt[(int64_t)j * P1 + r + 1];
            le = last ? scnt[j] : mycle[(int64_t)j * P1 + r + 1];
        }
        const long long supply = le - lt;
        const long long sp = supply > 0 ? supply : 0;
        const long long incl = wave_incl_scan_ll(sp);
        long long rem = deficit - (before + incl - sp);
        rem = rem > 0 ? rem : 0;
        if (j < P) {
            if (rem > 0 && supply < 0) atomicOr(error_flag, 4u);
            myC[(int64_t)j * P1 + r + 1] = lt + (sp < rem ? sp : rem);
        }
        before += __shfl(incl, 63, 64);
    }
    if (r == 0) {
        for (int j = lane; j < P; j += 64) {
            myC[(int64_t)j * P1] = 0;
            myclt[(int64_t)j * P1] = 0;
            mycle[(int64_t)j * P1] = 0;
            myclt[(int64_t)j * P1 + P] = scnt[j];
            mycle[(int64_t)j * P1 + P] = scnt[j];
        }
    }
}

// One wave per receiving rank r, lane j for senders j, j + 64, ...: send
// counts, receive displacements (the prefix of the counts over the senders)
// and the run table (receiver-major, i.e. in output order)
__global__ void runs_kernel(int P, const int64_t* __restrict__ C, const int64_t* __restrict__ sbeg,
                            const int64_t* __restrict__ myC, int64_t* __restrict__ send,
                            int64_t* __restrict__ run_dst, int64_t* __restrict__ run_src,
                            int* __restrict__ rnz, int* __restrict__ snz, uint32_t* __restrict__ error_flag) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= P) return;  // uniform per wave
    const int P1 = P + 1;
    long long displ = 0;  // counts of the senders before this chunk
    int nz = 0;           // non-empty runs of this receiver
    for (int j0 = 0; j0 < P; j0 += 64) {
        const int j = j0 + lane;
        int64_t a = 0, b = 0;
        if (j < P) {
            a = myC[(int64_t)j * P1 + r];
            b = myC[(int64_t)j * P1 + r + 1];
        }
        const int64_t cnt = b - a;
        const long long cp = cnt > 0 ? cnt : 0;
        const long long incl = wave_incl_scan_ll(cp);
        if (j < P) {
            if (cnt < 0) atomicOr(error_flag, 8u);
            send[(int64_t)j * P + r] = cnt;
            run_dst[(int64_t)r * P + j] = C[r] + displ + (incl - cp);
            run_src[(int64_t)r * P + j] = sbeg[j] + a;
            if (cnt > 0) atomicAdd(&snz[j], 1);
        }
        nz += __popcll(__ballot(j < P && cnt > 0));
        displ += __shfl(incl, 63, 64);
    }
    if (lane == 0) rnz[r] = nz;
    if (lane == 0 && displ != C[r + 1] - C[r]) atomicOr(error_flag, 16u);
}

// The exchange's choice (MPIU_Alltoallv mp-mpiu.c:125-136): dense when some
// sender has more than 128 non-empty peers (AUTO), or when asked; sparse
// otherwise, if the non-empty runs fit the compact table.  One block: the
// prefix of the receivers' non-empty runs (roff, P + 1 entries) and sel.
constexpr int kSparseVotePeers = 128;
__global__ __launch_bounds__(1024) void runs_vote_kernel(int P, const int* __restrict__ rnz,
                                                         const int* __restrict__ snz, int* __restrict__ roff,
                                                         int* __restrict__ sel, int policy, int cap) {
    __shared__ int s_w[16];
    __shared__ int s_dense;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_dense = 0;
    __syncthreads();
    int carry = 0;  // runs of the receivers before this chunk
    for (int r0 = 0; r0 < P; r0 += 1024) {
        const int r = r0 + tid;
        const int v = r < P ? rnz[r] : 0;
        if (r < P && snz[r] > kSparseVotePeers) s_dense = 1;  // r as a sender
        int x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        int pre = carry + x - v;
        for (int w = 0; w < wave; w++) pre += s_w[w];
        if (r < P) roff[r] = pre;
        int tot = 0;
        for (int w = 0; w < 16; w++) tot += s_w[w];
        __syncthreads();
        carry += tot;
    }
    if (tid == 0) {
        roff[P] = carry;
        const bool dense = policy == kExchDense || (policy == kExchAuto && s_dense) || carry > cap;
        sel[0] = dense ? 0 : 1;
        sel[1] = carry;
    }
}

// the compact table: a wave per receiver copies its non-empty runs (when
// the sparse table was chosen)
__global__ void runs_compact_kernel(int P, const int64_t* __restrict__ send, const int64_t* __restrict__ run_dst,
                                    const int64_t* __restrict__ run_src, const int* __restrict__ roff,
                                    const int* __restrict__ sel, int64_t* __restrict__ dst_c,
                                    int64_t* __restrict__ src_c) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= P || sel[0] == 0) return;  // uniform per wave
    int o = roff[r];
    for (int j0 = 0; j0 < P; j0 += 64) {
        const int j = j0 + lane;
        const bool nz = j < P && send[(int64_t)j * P + r] > 0;
        const uint64_t m = __ballot(nz);
        if (nz) {
            const int q = o + (int)__popcll(m & ((1ull << lane) - 1ull));
            dst_c[q] = run_dst[(int64_t)r * P + j];
            src_c[q] = run_src[(int64_t)r * P + j];
        }
        o += (int)__popcll(m);
    }
}

constexpr int kExchChunk = 4096;
constexpr int kExchLdsRuns = 512;

template <int RW>
__global__ __launch_bounds__(kBlock) void exchange_kernel(const uint32_t* __restrict__ src,
                                                          uint32_t* __restrict__ dst, int64_t ntot, RunTable rt,
                                                          uint32_t* __restrict__ error_flag) {
    __shared__ int64_t s_rd[kExchLdsRuns], s_rs[kExchLdsRuns];
    __shared__ int s_a, s_n;
    const int tid = threadIdx.x;
    const bool sparse = rt.sel && rt.sel[0] != 0;
    const int64_t* __restrict__ run_dst = sparse ? rt.dst_c : rt.dst;
    const int64_t* __restrict__ run_src = sparse ? rt.src_c : rt.src;
    const int nruns = sparse ? rt.sel[1] : rt.ndense;
    const int64_t o0 = (int64_t)blockIdx.x * kExchChunk;
    const int64_t o1 = min(ntot, o0 + (int64_t)kExchChunk);
    if (tid == 0) {
        const int a = upper_index<int64_t>(run_dst, nruns, o0);
        const int b = upper_index<int64_t>(run_dst, nruns, o1 - 1);
        s_a = a;
        s_n = b - a + 1;
    }
    __syncthreads();
    const int a = s_a, n = s_n;
    const bool local = n <= kExchLdsRuns;
    if (local) {
        for (int q = tid; q < n; q += kBlock) {
            s_rd[q] = run_dst[a + q];
            s_rs[q] = run_src[a + q];
        }
    }
    __syncthreads();
    // U records per thread at a time: sources found, then all U loads in
    // flight before the first store
    constexpr int U = 4;
    for (int64_t oo = o0 + tid; oo < o1; oo += U * kBlock) {
        int64_t sidx[U];
        uint32_t r[U][RW];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t o = oo + (int64_t)u * kBlock;
            sidx[u] = -1;
            if (o < o1) {
                if (local) {
                    const int q = upper_index<int64_t>(s_rd, n, o);
                    sidx[u] = s_rs[q] + (o - s_rd[q]);
                } else {
                    const int q = upper_index<int64_t>(run_dst, nruns, o);
                    sidx[u] = run_src[q] + (o - run_dst[q]);
                }
                if (sidx[u] < 0 || sidx[u] >= ntot) {  // only reachable with an inconsistent layout
                    atomicOr(error_flag, 32u);
                    sidx[u] = -1;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (sidx[u] >= 0) RecIO<RW>::load(src, sidx[u], r[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (sidx[u] >= 0) RecIO<RW>::store(dst, oo + (int64_t)u * kBlock, r[u]);
    }
}

// Exchange fused with the SecondSort's upfront histogram: each block also
// counts the digits of the records it moves for every digit place of its
// receiving rank (records are already in their receiver's segment after the
// move).  hist is [P][npass][256], cleared beforehand.
template <int RW>
__global__ __launch_bounds__(kBlock) void exchange_hist_kernel(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int64_t ntot, RunTable rt,
    uint32_t* __restrict__ error_flag, const int64_t* __restrict__ C, int P, int npass,
    PassTable pt, int b0, uint32_t* __restrict__ hist) {
    __shared__ int64_t s_rd[kExchLdsRuns], s_rs[kExchLdsRuns];
    extern __shared__ uint32_t s_hist[];  // npass * kBins counters
    __shared__ int s_a, s_n, s_r0, s_r1;
    const int tid = threadIdx.x;
    const bool sparse = rt.sel && rt.sel[0] != 0;
    const int64_t* __restrict__ run_dst = sparse ? rt.dst_c : rt.dst;
    const int64_t* __restrict__ run_src = sparse ? rt.src_c : rt.src;
    const int nruns = sparse ? rt.sel[1] : rt.ndense;
    const int nb = npass * kBins;
    const int64_t o0 = (int64_t)blockIdx.x * kExchChunk;
    const int64_t o1 = min(ntot, o0 + (int64_t)kExchChunk);
    if (tid == 0) {
        const int a = upper_index<int64_t>(run_dst, nruns, o0);
        const int b = upper_index<int64_t>(run_dst, nruns, o1 - 1);
        s_a = a;
        s_n = b - a + 1;
        s_r0 = upper_index<int64_t>(C, P, o0);
        s_r1 = upper_index<int64_t>(C, P, o1 - 1);
    }
    for (int i = tid; i < nb; i += kBlock) s_hist[i] = 0;
    __syncthreads();
    const int a = s_a, n = s_n, r0 = s_r0;
    const bool one_receiver = (r0 == s_r1);
    const bool local = n <= kExchLdsRuns;
    if (local) {
        for (int q = tid; q < n; q += kBlock) {
            s_rd[q] = run_dst[a + q];
            s_rs[q] = run_src[a + q];
        }
    }
    __syncthreads();
    // U records per thread at a time: sources found, then all U loads in
    // flight before the first store
    constexpr int U = 4;
    for (int64_t oo = o0 + tid; oo < o1; oo += U * kBlock) {
        int64_t sidx[U];
        uint32_t r[U][RW];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t o = oo + (int64_t)u * kBlock;
            sidx[u] = -1;
            if (o < o1) {
                if (local) {
                    const int q = upper_index<int64_t>(s_rd, n, o);
                    sidx[u] = s_rs[q] + (o - s_rd[q]);
                } else {
                    const int q = upper_index<int64_t>(run_dst, nruns, o);
                    sidx[u] = run_src[q] + (o - run_dst[q]);
                }
                if (sidx[u] < 0 || sidx[u] >= ntot) {
                    atomicOr(error_flag, 32u);
                    sidx[u] = -1;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (sidx[u] >= 0) RecIO<RW>::load(src, sidx[u], r[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (sidx[u] < 0) continue;
            const int64_t o = oo + (int64_t)u * kBlock;
            RecIO<RW>::store(dst, o, r[u]);
            if (one_receiver) {
                count_digits<RW>(r[u], npass, pt, b0, s_hist);
            } else {
                const int rr = upper_index<int64_t>(C, P, o);
                count_digits<RW>(r[u], npass, pt, b0, hist + (int64_t)rr * nb);
            }
        }
    }
    if (one_receiver) {
        __syncthreads();
        uint32_t* h = hist + (int64_t)r0 * nb;
        for (int i = tid; i < nb; i += kBlock) {
            const uint32_t c = s_hist[i];
            if (c) atomicAdd(&h[i], c);
        }
    }
}

// ---------------------------------------------------------------------------
// general record path: pack {key bytes, pad, index} sort records from the
// caller's records, and gather the caller's records by index at the end.
// The index is the record's position: one word (u32) in the last word of the
// sort record, or two words (u64, little-endian) in the last two when there
// are 2^32 records or more (iwords = 1 or 2).
// vec: records of a multiple of 16 bytes, 16-byte aligned (read as 16-byte
// vectors for the checksum; the key words come from the same lines again);
// kwords: key offset and length are whole words.
__global__ __launch_bounds__(kBlock) void pack_kernel(const uint8_t* __restrict__ in, int64_t n,
                                                      int64_t elsize, int64_t key_off, int key_bytes,
                                                      uint32_t* __restrict__ out, int rw, int iwords,
                                                      long long* __restrict__ checksum, int vec, int kwords) {
    __shared__ long long s_sum[kWaves];
    long long csum = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const uint8_t* rec = in + i * elsize;
        uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (checksum) {
            if (vec) {
                const uint4* v = reinterpret_cast<const uint4*>(rec);
                for (int64_t q = 0; q < (elsize >> 4); q++) {
                    const uint4 x = v[q];
                    csum += sbyte_sum(x.x) + sbyte_sum(x.y) + sbyte_sum(x.z) + sbyte_sum(x.w);
                }
            } else {
                for (int64_t b = 0; b < elsize; b++) csum += (signed char)rec[b];
            }
        }
        if (kwords) {
            const uint32_t* kw = reinterpret_cast<const uint32_t*>(rec + key_off);
            for (int q = 0; q < (key_bytes >> 2); q++) w[q] = kw[q];
        } else {
            for (int b = 0; b < key_bytes; b++) w[b >> 2] |= (uint32_t)rec[key_off + b] << ((b & 3) * 8);
        }
        if (iwords == 2) {
            w[rw - 2] = (uint32_t)i;
            w[rw - 1] = (uint32_t)((uint64_t)i >> 32);
        } else {
            w[rw - 1] = (uint32_t)i;
        }
        if (rw == 4) {
            reinterpret_cast<uint4*>(out)[i] = make_uint4(w[0], w[1], w[2], w[3]);
        } else if (rw == 2) {
            reinterpret_cast<uint2*>(out)[i] = make_uint2(w[0], w[1]);
        } else {
            for (int q = 0; q < rw; q++) out[i * rw + q] = w[q];
        }
    }
    if (checksum) {
        csum = wave_sum_ll(csum);
        if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = csum;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long a = 0;
            for (int q = 0; q < kWaves; q++) a += s_sum[q];
            atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
        }
    }
}

__global__ __launch_bounds__(kBlock) void gather_kernel(const uint8_t* __restrict__ in, int64_t n,
                                                        int64_t elsize, const uint32_t* __restrict__ srec,
                                                        int rw, int iwords, uint8_t* __restrict__ out,
                                                        long long* __restrict__ checksum,
                                                        uint32_t* __restrict__ error_flag) {
    __shared__ long long s_sum[kWaves];
    long long csum = 0;
    const bool words = (elsize & 3) == 0 && ((uintptr_t)in & 3) == 0 && ((uintptr_t)out & 3) == 0;
    const bool vec = (elsize & 15) == 0 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0;
    for (int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x; o < n;
         o += (int64_t)gridDim.x * kBlock) {
        const int64_t i = iwords == 2 ? (int64_t)(((uint64_t)srec[o * rw + rw - 1] << 32) | srec[o * rw + rw - 2])
                                      : (int64_t)srec[o * rw + rw - 1];
        if (i < 0 || i >= n) {  // only reachable with corrupted sort records
            if (error_flag) atomicOr(error_flag, 64u);
            continue;
        }
        if (vec) {
            const uint4* s = reinterpret_cast<const uint4*>(in + i * elsize);
            uint4* d = reinterpret_cast<uint4*>(out + o * elsize);
            for (int64_t q = 0; q < (elsize >> 4); q++) {
                const uint4 v = s[q];
                d[q] = v;
                if (checksum) csum += sbyte_sum(v.x) + sbyte_sum(v.y) + sbyte_sum(v.z) + sbyte_sum(v.w);
            }
        } else if (words) {
            const uint32_t* s = reinterpret_cast<const uint32_t*>(in + i * elsize);
            uint32_t* d = reinterpret_cast<uint32_t*>(out + o * elsize);
            for (int64_t q = 0; q < (elsize >> 2); q++) {
                const uint32_t v = s[q];
                d[q] = v;
                if (checksum) csum += sbyte_sum(v);
            }
        } else {
            for (int64_t b = 0; b < elsize; b++) {
                const uint8_t v = in[i * elsize + b];
                out[o * elsize + b] = v;
                if (checksum) csum += (signed char)v;
            }
        }
    }
    if (checksum) {
        csum = wave_sum_ll(csum);
        if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = csum;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long a = 0;
            for (int q = 0; q < kWaves; q++) a += s_sum[q];
            atomicAdd((unsigned long long*)checksum, (unsigned long long)a);
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
int launch_bisect(const void* recs, int rw, const KeySpec& ks, const int64_t* sbeg,
                  const int64_t* scnt, int P, const int64_t* C, int64_t* win, int64_t* myclt,
                  int64_t* mycle, uint64_t* split, uint64_t* pminmax, int* rounds, uint32_t* err,
                  hipStream_t st, const ModelDev* model, int* exact, bool perturb) {
    if (P < 2) return 0;
    const int blocks = (P - 1 + kWaves - 1) / kWaves;
    ModelDev md{};
    if (model) md = *model;
    hipLaunchKernelGGL(bisect_kernel, dim3(blocks), dim3(kBlock), 0, st, (const uint32_t*)recs, rw,
                       ks, sbeg, scnt, P, C, win, myclt, mycle, split, pminmax, rounds, err, md,
                       model ? (perturb ? 2 : 1) : 0, exact);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_layout(int P, const int64_t* C, const int64_t* scnt, const int64_t* sbeg, int64_t* myclt,
                  int64_t* mycle, int64_t* myC, int64_t* send, const ExchRuns& ex, uint32_t* err, hipStream_t st) {
    const int blocks = (P + kWaves - 1) / kWaves;  // a wave per receiving rank
    hipLaunchKernelGGL(layout_kernel, dim3(blocks), dim3(kBlock), 0, st, P, C, scnt, myclt, mycle,
                       myC, ex.snz, err);
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(runs_kernel, dim3(blocks), dim3(kBlock), 0, st, P, C, sbeg, myC, send, ex.dst, ex.src,
                       ex.rnz, ex.snz, err);
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(runs_vote_kernel, dim3(1), dim3(1024), 0, st, P, ex.rnz, ex.snz, ex.roff, ex.sel, ex.policy,
                       ex.cap);
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(runs_compact_kernel, dim3(blocks), dim3(kBlock), 0, st, P, send, ex.dst, ex.src, ex.roff,
                       ex.sel, ex.dst_c, ex.src_c);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_exchange(const void* src, void* dst, int64_t ntot, int rw, const RunTable& rt, uint32_t* err,
                    hipStream_t st) {
    if (ntot <= 0) return 0;
    const int64_t blocks = (ntot + kExchChunk - 1) / kExchChunk;
    dim3 g((unsigned)blocks), b(kBlock);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    switch (rw) {
        case 1: hipLaunchKernelGGL(exchange_kernel<1>, g, b, 0, st, s, d, ntot, rt, err); break;
        case 2: hipLaunchKernelGGL(exchange_kernel<2>, g, b, 0, st, s, d, ntot, rt, err); break;
        case 4: hipLaunchKernelGGL(exchange_kernel<4>, g, b, 0, st, s, d, ntot, rt, err); break;
        case 8: hipLaunchKernelGGL(exchange_kernel<8>, g, b, 0, st, s, d, ntot, rt, err); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_exchange_hist(const void* src, void* dst, int64_t ntot, int rw, const RunTable& rt, uint32_t* err,
                         const int64_t* C, int P, const KeySpec& ks, uint32_t* hist, hipStream_t st) {
    if (ntot <= 0) return 0;
    const int64_t blocks = (ntot + kExchChunk - 1) / kExchChunk;
    dim3 g((unsigned)blocks), b(kBlock);
    const uint32_t* s = (const uint32_t*)src;
    uint32_t* d = (uint32_t*)dst;
    const PassTable pt = make_pass_table(ks);
    const int np = ks.key_bytes;
    const int b0 = consecutive_key_bytes(pt, np);
    const size_t lds = (size_t)np * kBins * sizeof(uint32_t);
    switch (rw) {
        case 1: hipLaunchKernelGGL(exchange_hist_kernel<1>, g, b, lds, st, s, d, ntot, rt, err, C, P, np, pt, b0, hist); break;
        case 2: hipLaunchKernelGGL(exchange_hist_kernel<2>, g, b, lds, st, s, d, ntot, rt, err, C, P, np, pt, b0, hist); break;
        case 4: hipLaunchKernelGGL(exchange_hist_kernel<4>, g, b, lds, st, s, d, ntot, rt, err, C, P, np, pt, b0, hist); break;
        case 8: hipLaunchKernelGGL(exchange_hist_kernel<8>, g, b, lds, st, s, d, ntot, rt, err, C, P, np, pt, b0, hist); break;
        default: return -1;
    }
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_pack(const void* in, int64_t n, int64_t elsize, int64_t key_off, int key_bytes,
                void* out, int rw, int iwords, long long* checksum, hipStream_t st) {
    if (n <= 0) return 0;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    const int vec = (elsize % 16 == 0 && (uintptr_t)in % 16 == 0) ? 1 : 0;
    const int kwords = (key_off % 4 == 0 && key_bytes % 4 == 0 && elsize % 4 == 0 && (uintptr_t)in % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, (const uint8_t*)in,
                       n, elsize, key_off, key_bytes, (uint32_t*)out, rw, iwords, checksum, vec, kwords);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_gather(const void* in, int64_t n, int64_t elsize, const void* srec, int rw, int iwords, void* out,
                  long long* checksum, uint32_t* err, hipStream_t st) {
    if (n <= 0) return 0;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, st, (const uint8_t*)in,
                       n, elsize, (const uint32_t*)srec, rw, iwords, (uint8_t*)out, checksum, err);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

// The corruption fault hook: one byte of a sort record flipped (bit 0), or,
// without a record to change, 1 added to the recorded input sum; either way
// the output sum no longer matches.
__global__ void fault_corrupt_kernel(uint8_t* __restrict__ byte, long long* __restrict__ checksum_in) {
    if (byte) *byte ^= 1u;
    else if (checksum_in) *checksum_in += 1;
}

int launch_fault_corrupt(void* byte, long long* checksum_in, hipStream_t st) {
    hipLaunchKernelGGL(fault_corrupt_kernel, dim3(1), dim3(1), 0, st, (uint8_t*)byte, checksum_in);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

}  // namespace mpsg
