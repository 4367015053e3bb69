// model.hip -- the learned key-CDF model of the learned-splitter path.
//
// F(key) in [0, 1] approximates the fraction of records with radix <= key.
// It is a two-level learned index:
//   root   a monotone MLP 1 -> 32 -> 32 -> 1 on x = (key - kmin) / (kmax - kmin)
//          or, for keys spread over many orders of magnitude,
//          x = log2(1 + key - kmin) / log2(1 + kmax - kmin) (both monotone):
//              h1 = tanh(w1 * x + b1), h2 = tanh(W2 h1 + b2), r = w3 . h2 + b3
//          with w1, W2, w3 = softplus(raw) >= 0 and monotone activations, so r
//          is non-decreasing in x.  Trained here on the GPU (forward/backward
//          kernels below, Adam) on sampled keys with their exact ranks.
//   leaves L linear pieces; leaf = clamp(floor(r * L), 0, L-1) and
//          F = clamp(a_leaf * x + b_leaf, lo_leaf, hi_leaf) with a_leaf >= 0 and
//          [lo, hi] the range of sampled ranks routed to the leaf.  Samples are
//          sorted and routing is monotone, so hi_leaf <= lo_leaf+1 and F is
//          non-decreasing across leaves.
// Everything is FP32 on the vector ALUs.  The model never decides an output
// position: it only proposes splitter values (checked and repaired with exact
// counts) and bucket boundaries (integer keys, compared exactly).
#include "device.h"
#include "model.h"
#include "model_dev.h"

namespace mpsg {

__device__ __forceinline__ float softplus_f(float v) {
    return v > 20.f ? v : log1pf(expf(v));
}
__device__ __forceinline__ float sigmoid_f(float v) { return 1.f / (1.f + expf(-v)); }

// effective (non-negative) weights from raw parameters
__global__ void mlp_effective_kernel(const float* __restrict__ raw, float* __restrict__ eff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float v = raw[i];
    const bool positive = (i >= kW1 && i < kW1 + kH) || (i >= kW2 && i < kW2 + kH * kH) ||
                          (i >= kW3 && i < kW3 + kH);
    eff[i] = positive ? softplus_f(v) : v;
}

// Loss = mean over the batch of (r(x_i) - y_i)^2.  Block b computes the
// gradient of its samples and writes it to partial[b][kMlpParams] (no float
// atomics: the reduction order is fixed, so training is deterministic).
constexpr int kGradBlock = 256;
__global__ __launch_bounds__(kGradBlock) void mlp_grad_kernel(const float* __restrict__ x,
                                                              const float* __restrict__ y, int64_t n,
                                                              int64_t stride, int64_t offset,
                                                              const float* __restrict__ raw,
                                                              const float* __restrict__ eff_g,
                                                              float* __restrict__ partial,
                                                              float* __restrict__ loss_partial) {
    __shared__ float e[kMlpParams];
    __shared__ float s_h1[kGradBlock][kH + 1];
    __shared__ float s_dz2[kGradBlock][kH + 1];
    __shared__ float s_red[(kGradBlock / 64) * (3 * kH + 2)];
    const int tid = threadIdx.x;
    for (int i = tid; i < kMlpParams; i += kGradBlock) e[i] = eff_g[i];
    __syncthreads();
    const int64_t si = offset + ((int64_t)blockIdx.x * kGradBlock + tid) * stride;
    const bool valid = si < n;
    const float inv_n = 1.f / (float)((n - offset + stride - 1) / stride);
    float h1[kH], h2[kH];
    float xv = valid ? x[si] : 0.f;
    float r = mlp_forward(e, xv, h1, h2);
    const float dy = valid ? 2.f * (r - y[si]) * inv_n : 0.f;
    const float lv = valid ? (r - y[si]) * (r - y[si]) * inv_n : 0.f;
    // backward
    float dz1[kH];
#pragma unroll
    for (int j = 0; j < kH; j++) dz1[j] = 0.f;
#pragma unroll 4
    for (int k = 0; k < kH; k++) {
        const float dz2 = dy * e[kW3 + k] * (1.f - h2[k] * h2[k]);
        s_dz2[tid][k] = dz2;
#pragma unroll
        for (int j = 0; j < kH; j++) dz1[j] += dz2 * e[kW2 + k * kH + j];
    }
#pragma unroll
    for (int j = 0; j < kH; j++) {
        dz1[j] *= (1.f - h1[j] * h1[j]);
        s_h1[tid][j] = h1[j];
    }
    // per-sample small gradients (w1, b1, w3: 3 x 32, b3, loss): wave
    // shuffle sums, then a fixed-order sum over the block's waves
    float* out = partial + (int64_t)blockIdx.x * kMlpParams;
    const int lane = tid & 63, wv = tid >> 6;
    constexpr int kSmall = 3 * kH + 2;
    for (int q = 0; q < kSmall; q++) {
        float v;
        if (q < kH) v = dz1[q] * xv;
        else if (q < 2 * kH) v = dz1[q - kH];
        else if (q < 3 * kH) v = dy * h2[q - 2 * kH];
        else if (q == 3 * kH) v = dy;
        else v = lv;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) s_red[wv * kSmall + q] = v;
    }
    __syncthreads();
    if (tid < kSmall) {
        float acc = 0.f;
        for (int w = 0; w < kGradBlock / 64; w++) acc += s_red[w * kSmall + tid];
        if (tid < kH) out[kW1 + tid] = acc * sigmoid_f(raw[kW1 + tid]);
        else if (tid < 2 * kH) out[kB1 + tid - kH] = acc;
        else if (tid < 3 * kH) out[kW3 + tid - 2 * kH] = acc * sigmoid_f(raw[kW3 + tid - 2 * kH]);
        else if (tid == 3 * kH) out[kB3] = acc;
        else loss_partial[blockIdx.x] = acc;
    }
    // b2[k] = sum dz2[:,k]; W2[k][j] = sum dz2[:,k] * h1[:,j]  (1024 + 32 outputs)
    __syncthreads();
    for (int o = tid; o < kH * kH + kH; o += kGradBlock) {
        float acc = 0.f;
        if (o < kH * kH) {
            const int k = o / kH, j = o % kH;
            for (int s = 0; s < kGradBlock; s++) acc += s_dz2[s][k] * s_h1[s][j];
            out[kW2 + o] = acc * sigmoid_f(raw[kW2 + o]);
        } else {
            const int k = o - kH * kH;
            for (int s = 0; s < kGradBlock; s++) acc += s_dz2[s][k];
            out[kB2 + k] = acc;
        }
    }
}

// sum the per-block partial gradients (fixed order) and apply one Adam step
__global__ void mlp_adam_kernel(float* __restrict__ raw, const float* __restrict__ partial, int nblocks,
                                float* __restrict__ m, float* __restrict__ v, float lr, float beta1,
                                float beta2, float eps, int step, float* __restrict__ grad_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; b++) g += partial[(int64_t)b * kMlpParams + i];
    if (grad_out) grad_out[i] = g;
    if (lr <= 0.f) return;  // gradient only
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(beta1, (float)step));
    const float vh = vi / (1.f - powf(beta2, (float)step));
    raw[i] -= lr * mh / (sqrtf(vh) + eps);
}

// root prediction of every sample -> leaf index, made monotone over the sorted
// samples by a running maximum (done by the caller's scan kernel below)
__global__ void mlp_leaf_kernel(const float* __restrict__ x, int64_t n, const float* __restrict__ eff_g,
                                int L, int linear_root, int* __restrict__ leaf) {
    __shared__ float e[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) e[i] = eff_g[i];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = linear_root ? x[i] : mlp_forward_only(e, x[i]);
    int l = (int)floorf(r * (float)L);
    l = l < 0 ? 0 : (l >= L ? L - 1 : l);
    leaf[i] = l;
}

// single-block: running max of leaf[] (monotone routing) and leaf start offsets
__global__ void leaf_starts_kernel(int* __restrict__ leaf, int64_t n, int L, int64_t* __restrict__ start) {
    // serial over samples in chunks: a single block of 1024 threads does a
    // blocked running max then fills starts
    __shared__ int s_max[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t per = (n + nt - 1) / nt;
    const int64_t a = tid * per, b = min(n, a + per);
    int m = -1;
    for (int64_t i = a; i < b; i++) m = max(m, leaf[i]);
    s_max[tid] = m;
    __syncthreads();
    if (tid == 0) {
        int run = -1;
        for (int t = 0; t < nt; t++) {
            const int v = s_max[t];
            s_max[t] = run;  // exclusive running max
            run = max(run, v);
        }
    }
    __syncthreads();
    m = s_max[tid];
    for (int64_t i = a; i < b; i++) {
        m = max(m, leaf[i]);
        leaf[i] = m;
    }
    __syncthreads();
    // start[l] = first sample with leaf >= l; start[L] = n
    for (int l = tid; l <= L; l += nt) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (leaf[mid] < l) lo = mid + 1; else hi = mid;
        }
        start[l] = lo;
    }
}

// least-squares line per leaf, slope clamped to >= 0, output range [lo, hi]
__global__ void leaf_fit_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                const int64_t* __restrict__ start, int L, float* __restrict__ leaves,
                                float ylast_default) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const int64_t a = start[l], b = start[l + 1];
    float* o = leaves + 4 * (int64_t)l;
    if (b <= a) {
        // empty leaf: constant at the last rank before it
        const float c = a > 0 ? y[a - 1] : 0.f;
        o[0] = 0.f; o[1] = c; o[2] = c; o[3] = c;
        return;
    }
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    const double cnt = (double)(b - a);
    const double x0 = x[a];
    for (int64_t i = a; i < b; i++) {
        const double dx = (double)x[i] - x0, yy = y[i];
        sx += dx; sy += yy; sxx += dx * dx; sxy += dx * yy;
    }
    const double den = cnt * sxx - sx * sx;
    double slope = den > 0 ? (cnt * sxy - sx * sy) / den : 0.0;
    if (!(slope >= 0)) slope = 0.0;
    const double icpt = (sy - slope * sx) / cnt - slope * x0;
    o[0] = (float)slope;
    o[1] = (float)icpt;
    o[2] = y[a];
    o[3] = y[b - 1];
    (void)ylast_default;
}

__global__ void model_eval_kernel(ModelDev m, const uint64_t* __restrict__ keys, int64_t n,
                                  float* __restrict__ out) {
    __shared__ float e[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) e[i] = m.eff[i];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = model_eval(e, m, keys[i]);
}

// leaf of x under the root (as model_eval routes it)
__device__ __forceinline__ int leaf_of_x(const float* __restrict__ e, const ModelDev& m, float x) {
    const float r = m.linear_root ? x : mlp_forward_only(e, x);
    int l = (int)floorf(r * (float)m.L);
    return l < 0 ? 0 : (l >= m.L ? m.L - 1 : l);
}

// xstart[l]: the smallest float x in [0, 1] with leaf_of_x(x) >= l, by
// bisection on the bit pattern (non-negative floats order like their bits)
__global__ void leaf_xstart_kernel(ModelDev m, float* __restrict__ xstart) {
    __shared__ float e[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) e[i] = m.eff[i];
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > m.L) return;
    if (l == 0) { xstart[0] = 0.f; return; }
    if (l == m.L) { xstart[l] = 2.f; return; }
    uint32_t lo = 0, hi = 0x3f800000u;  // [0, 1.0]
    if (leaf_of_x(e, m, 1.0f) < l) { xstart[l] = 2.f; return; }
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (leaf_of_x(e, m, __uint_as_float(mid)) >= l) hi = mid; else lo = mid + 1;
    }
    xstart[l] = __uint_as_float(lo);
}

// Boundary tables: table t has nb boundaries for quantile targets
// q = q0[t] + (q1[t] - q0[t]) * (i + 1) / (nb + 1).  With nedge[t] = E > 0
// the first E targets split [0, q0) evenly, the last E split (q1, 1], and
// the nb - 2E between them split [q0, q1]: a table for a segment whose keys
// mostly lie in one part of the distribution that still spreads the few
// keys outside it.  Tables are made non-decreasing afterwards.  One thread
// per boundary.
__global__ void model_tables_kernel(ModelDev m, const float* __restrict__ q0, const float* __restrict__ q1,
                                    const int* __restrict__ nedge, int ntables, int nb,
                                    uint64_t* __restrict__ tables, const uint64_t* __restrict__ heavy,
                                    const float2* __restrict__ heavy_q, int nheavy) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)ntables * nb) return;
    const int t = (int)(g / nb), i = (int)(g % nb);
    const int E = nedge ? nedge[t] : 0;
    const float a = q0[t], b = q1[t];
    float q;
    if (E > 0 && i < E) {
        q = a * (float)(i + 1) / (float)(E + 1);
    } else if (E > 0 && i >= nb - E) {
        q = b + (1.f - b) * (float)(i - (nb - E) + 1) / (float)(E + 1);
    } else {
        q = a + (b - a) * (float)(i - E + 1) / (float)(nb - 2 * E + 1);
    }
    uint64_t k = model_inverse_leaves(m, q);
    if (nheavy > 0) {
        // exact at the heavy keys, whatever the model's error at their
        // jumps: q inside heavy key h's sampled range [qa, qb) gives h's
        // key; otherwise the key lies strictly between the heavy keys whose
        // ranges end before q and start after it
        int lo = 0, hi = nheavy;  // lo = #{heavy keys with qb <= q}
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (heavy_q[mid].y <= q) lo = mid + 1; else hi = mid;
        }
        if (lo < nheavy && heavy_q[lo].x <= q) {
            k = heavy[lo];
        } else {
            if (lo < nheavy && heavy[lo] > 0 && k >= heavy[lo]) k = heavy[lo] - 1;
            if (lo > 0 && k <= heavy[lo - 1] && heavy[lo - 1] < ~0ull) k = heavy[lo - 1] + 1;
        }
    }
    tables[g] = k;
}

// Quantile range of each segment's keys under the model, from 256 keys read
// at evenly spaced positions of the segment (unsorted records of elsize
// bytes, key_bytes-byte key at key_off, XOR key_xmask): the narrowest
// interval of F values that holds 246 of the 256 (so up to 10 keys from
// elsewhere in the distribution, on either side, do not widen it).  A
// segment whose interval covers less than half of [0, 1] (its keys come from
// one part of the distribution, as in nearly sorted input) gets [q0, q1] =
// that interval widened by 1/64 of its width on each side and nedge =
// kEdgeBounds boundaries on each side for the keys outside it; the others get
// [0, 1] and no edges.  One wave per segment.
constexpr int kEdgeBounds = 32;
__global__ __launch_bounds__(256) void seg_range_kernel(ModelDev m, const uint8_t* __restrict__ recs, int64_t elsize,
                                                        int64_t key_off, int key_bytes, uint64_t key_xmask,
                                                        const int64_t* __restrict__ begin,
                                                        const int64_t* __restrict__ count, float* __restrict__ q0,
                                                        float* __restrict__ q1, int* __restrict__ nedge) {
    __shared__ float s_v[256];  // the model's value at sample i
    __shared__ float s_f[256];  // the same values, sorted
    __shared__ float s_w[64];
    const int s = blockIdx.x, tid = threadIdx.x, lane = tid;
    const int64_t b = begin[s], c = count[s];
    // sample i = tid: the record at b + floor((i + 1/2) c / 256)
    float f = 0.f;
    if (c > 0) {
        int64_t pos = b + (int64_t)(((double)tid + 0.5) * (double)c / 256.0);
        pos = pos >= b + c ? b + c - 1 : pos;
        const uint8_t* p = recs + pos * elsize + key_off;
        uint64_t k = 0;
        for (int q = 0; q < key_bytes; q++) k |= (uint64_t)p[q] << (8 * q);
        f = model_eval(m.eff, m, k ^ key_xmask);
    }
    s_v[tid] = f;
    __syncthreads();
    // rank of each value among the 256 (ties by sample index)
    int rank = 0;
    for (int i = 0; i < 256; i++) {
        const float v = s_v[i];
        rank += (v < f || (v == f && i < tid)) ? 1 : 0;
    }
    s_f[rank] = f;
    __syncthreads();
    // windows of 246 consecutive sorted values starting at 0 .. 10
    constexpr int kKeep = 246;
    if (lane < 64) s_w[lane] = lane <= 256 - kKeep ? s_f[lane + kKeep - 1] - s_f[lane] : 2.f;
    __syncthreads();
    if (lane == 0) {
        int best = 0;
        for (int q = 1; q <= 256 - kKeep; q++)
            if (s_w[q] < s_w[best]) best = q;
        const float lo = s_f[best], hi = s_f[best + kKeep - 1], w = hi - lo;
        if (c < 4096 || !(w <= 0.5f)) {
            q0[s] = 0.f;
            q1[s] = 1.f;
            nedge[s] = 0;
        } else {
            const float pad = w / 64.f + 1e-5f;
            q0[s] = fmaxf(0.f, lo - pad);
            q1[s] = fminf(1.f, hi + pad);
            nedge[s] = kEdgeBounds;
        }
    }
}

// per table (one block each, nb <= 4096): running max so the table is sorted
// even if FP32 rounding made F locally non-monotone; then every heavy key v
// (ascending) that weighs at least one of the table's buckets (sampled
// weight qb - qa >= (q1 - q0) / (nb - 2 nedge)) and lies inside its span
// [t[0], t[nb - 1]] or within (q1 - q0) / 4 of its quantile range (the data
// sorted may hold a little more of v than the training sample, spilling
// into this table's first or last bucket) gets the boundary pair {v - 1, v},
// i.e. a bucket holding v alone.
// Overwriting the two boundaries around v's position keeps the table sorted;
// a pair that would overwrite the previous heavy key's pair moves just after
// it, or is dropped when the table has no room left.
constexpr int kFixHeavy = 16;             // heavy keys per thread (kMaxHeavy / 1024)
constexpr int kFixList = 2048;            // heavy keys one table takes, at most
__global__ __launch_bounds__(1024) void tables_fix_kernel(uint64_t* __restrict__ tables, int ntables, int nb,
                                                          const uint64_t* __restrict__ heavy,
                                                          const float2* __restrict__ heavy_q, int nheavy,
                                                          const float* __restrict__ q0, const float* __restrict__ q1,
                                                          const int* __restrict__ nedge) {
    (void)ntables;  // one block per table
    constexpr int E = 4;  // entries per thread
    __shared__ uint64_t s_t[1024 * E];
    __shared__ uint64_t s_w[16];
    __shared__ uint64_t s_h[kFixList];  // this table's heavy keys, ascending
    __shared__ uint32_t s_c[16];
    __shared__ uint32_t s_nh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t* tb = tables + (int64_t)blockIdx.x * nb;
    // entries E tid .. E tid + E - 1: their max, then a running max over
    // threads
    uint64_t e[E];
#pragma unroll
    for (int q = 0; q < E; q++) e[q] = E * tid + q < nb ? tb[E * tid + q] : 0;
    uint64_t v = 0;
#pragma unroll
    for (int q = 0; q < E; q++) v = e[q] > v ? e[q] : v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t u = __shfl_up(v, off, 64);
        if (lane >= off && u > v) v = u;
    }
    if (lane == 63) s_w[wave] = v;
    __syncthreads();
    uint64_t pre = 0;
    for (int w = 0; w < wave; w++) pre = s_w[w] > pre ? s_w[w] : pre;
    // running max before this thread's entries
    uint64_t before = __shfl_up(v, 1, 64);
    if (lane == 0) before = 0;
    before = pre > before ? pre : before;
#pragma unroll
    for (int q = 0; q < E; q++) {
        before = before > e[q] ? before : e[q];
        if (E * tid + q < nb) s_t[E * tid + q] = before;
    }
    __syncthreads();
    {
        // this table's heavy keys, in key order: kFixHeavy consecutive ones
        // per thread, kept by a block prefix count
        const int t = blockIdx.x;
        const float qa = q0 ? q0[t] : 0.f, qb = q1 ? q1[t] : 1.f;
        const int ne = nedge ? nedge[t] : 0;
        const float wmin = (qb - qa) / (float)(nb - 2 * ne > 1 ? nb - 2 * ne : 1);
        const uint64_t tfirst = s_t[0], tlast = s_t[nb - 1];
        const float qtol = (qb - qa) * 0.25f;
        uint32_t keep = 0, cnt = 0;
#pragma unroll
        for (int q = 0; q < kFixHeavy; q++) {
            const int h = tid * kFixHeavy + q;
            if (h < nheavy) {
                const float2 hq = heavy_q[h];
                const uint64_t hv = heavy[h];
                const bool ok = hq.y - hq.x >= wmin &&
                                ((hv >= tfirst && hv <= tlast) || (hq.y >= qa - qtol && hq.x <= qb + qtol));
                keep |= ok ? 1u << q : 0u;
                cnt += ok ? 1u : 0u;
            }
        }
        uint32_t x = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_c[wave] = x;
        __syncthreads();
        uint32_t pre = x - cnt;
        for (int w = 0; w < wave; w++) pre += s_c[w];
        if (tid == 1023) s_nh = pre + cnt;
#pragma unroll
        for (int q = 0; q < kFixHeavy; q++)
            if ((keep >> q) & 1u) {
                if (pre < (uint32_t)kFixList) s_h[pre] = heavy[tid * kFixHeavy + q];
                pre++;
            }
    }
    __syncthreads();
    if (tid == 0 && nb >= 2) {
        int last = -1;  // slot of the previous heavy key
        const int nh = s_nh < (uint32_t)kFixList ? (int)s_nh : kFixList;
        for (int h = 0; h < nh; h++) {
            const uint64_t hv = s_h[h];
            if (hv == 0) {
                s_t[0] = 0;  // bucket 0 = keys <= 0
                last = 0;
                continue;
            }
            int a = 0, b = nb;  // a = first slot with s_t[a] >= hv
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (s_t[mid] < hv) a = mid + 1; else b = mid;
            }
            if (a < nb && s_t[a] == hv && a > 0 && s_t[a - 1] == hv - 1) {
                last = a;
                continue;
            }
            int j = a < 1 ? 1 : (a >= nb ? nb - 1 : a);
            if (j - 1 <= last) {
                // s_t[last] holds the previous heavy key u < hv
                j = (s_t[last] == hv - 1) ? last + 1 : last + 2;
                if (j >= nb) break;
            }
            s_t[j - 1] = hv - 1;
            s_t[j] = hv;
            last = j;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; q++)
        if (E * tid + q < nb) tb[E * tid + q] = s_t[E * tid + q];
}

// tables of consecutive quantile ranges, one after the other (one block per
// table after the first): entries below the largest entry of the earlier
// tables are raised to it, so that the tables read as one sorted sequence
// (two tables that both isolate a heavy key on their common edge would
// otherwise overlap by its pair)
__global__ __launch_bounds__(256) void tables_order_kernel(uint64_t* __restrict__ tables, int ntables, int nb) {
    (void)ntables;
    __shared__ uint64_t s_m;
    const int t = blockIdx.x + 1;
    if (threadIdx.x == 0) {
        uint64_t m = 0;
        for (int u = 0; u < t; u++) {
            const uint64_t e = tables[(int64_t)u * nb + nb - 1];  // each table is sorted
            m = e > m ? e : m;
        }
        s_m = m;
    }
    __syncthreads();
    const uint64_t m = s_m;
    uint64_t* tb = tables + (int64_t)t * nb;
    for (int i = threadIdx.x; i < nb; i += 256)
        if (tb[i] < m) tb[i] = m;
}

// training samples from a sorted record array: S stratified positions with
// exact ranks; x = normalised key, y = (rank + 1) / n
__global__ void model_sample_kernel(const uint8_t* __restrict__ recs, int64_t elsize, int64_t key_off,
                                    int key_bytes, uint64_t key_xmask, int64_t n, int64_t S, uint64_t seed,
                                    ModelDev m, uint64_t* __restrict__ skeys,
                                    float* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    uint64_t h = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    const double u = (double)(h >> 11) * 0x1.0p-53;
    int64_t pos = (int64_t)(((double)i + u) * (double)n / (double)S);
    if (pos >= n) pos = n - 1;
    const uint8_t* p = recs + pos * elsize + key_off;
    uint64_t k = 0;
    for (int b = 0; b < key_bytes; b++) k |= (uint64_t)p[b] << (8 * b);
    k ^= key_xmask;
    skeys[i] = k;
    x[i] = key_to_x(k, m);
    y[i] = (float)((double)(pos + 1) / (double)n);
}

// ---------------------------------------------------------------------------
// host launchers
static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_mlp_effective(const float* raw, float* eff, hipStream_t st) {
    hipLaunchKernelGGL(mlp_effective_kernel, dim3(nblk(kMlpParams, 256)), dim3(256), 0, st, raw, eff);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int mlp_grad_blocks(int64_t n, int64_t stride, int64_t offset) {
    const int64_t m = (n - offset + stride - 1) / stride;
    return (int)nblk(m, kGradBlock);
}

int launch_mlp_grad(const float* x, const float* y, int64_t n, int64_t stride, int64_t offset,
                    const float* raw, const float* eff, float* partial, float* loss_partial, hipStream_t st) {
    const int blocks = mlp_grad_blocks(n, stride, offset);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(kGradBlock), 0, st, x, y, n, stride, offset, raw,
                       eff, partial, loss_partial);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_mlp_adam(float* raw, const float* partial, int nblocks, float* m, float* v, float lr, int step,
                    float* grad_out, hipStream_t st) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(nblk(kMlpParams, 256)), dim3(256), 0, st, raw, partial, nblocks,
                       m, v, lr, 0.9f, 0.999f, 1e-8f, step, grad_out);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_leaves(const float* x, const float* y, int64_t S, const float* eff, int L, int linear_root,
                  int* leaf, int64_t* start, float* leaves, hipStream_t st) {
    hipLaunchKernelGGL(mlp_leaf_kernel, dim3(nblk(S, 256)), dim3(256), 0, st, x, S, eff, L, linear_root, leaf);
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(leaf_starts_kernel, dim3(1), dim3(1024), 0, st, leaf, S, L, start);
    MPSG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(leaf_fit_kernel, dim3(nblk(L, 256)), dim3(256), 0, st, x, y, start, L, leaves, 0.f);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_leaf_xstart(const float* eff, int L, int linear_root, float* xstart, hipStream_t st) {
    ModelDev m{};
    m.eff = eff;
    m.L = L;
    m.linear_root = linear_root;
    hipLaunchKernelGGL(leaf_xstart_kernel, dim3(nblk(L + 1, 128)), dim3(128), 0, st, m, xstart);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_model_eval(const ModelDev& m, const uint64_t* keys, int64_t n, float* out, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(model_eval_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, m, keys, n, out);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_model_tables(const ModelDev& m, uint64_t kmax, const float* q0, const float* q1, int ntables,
                        int nb, uint64_t* tables, const uint64_t* heavy, const float2* heavy_q, int nheavy,
                        hipStream_t st, const int* nedge, bool ordered) {
    const int64_t tot = (int64_t)ntables * nb;
    (void)kmax;
    hipLaunchKernelGGL(model_tables_kernel, dim3(nblk(tot, 128)), dim3(128), 0, st, m, q0, q1, nedge, ntables, nb,
                       tables, heavy, heavy_q, nheavy);
    MPSG_CHECK(hipGetLastError());
    if (nb > 4096 || nheavy > kMaxHeavy) return -1;
    hipLaunchKernelGGL(tables_fix_kernel, dim3(ntables), dim3(1024), 0, st, tables, ntables, nb, heavy, heavy_q,
                       nheavy, q0, q1, nedge);
    MPSG_CHECK(hipGetLastError());
    if (ordered && ntables > 1) {
        hipLaunchKernelGGL(tables_order_kernel, dim3(ntables - 1), dim3(256), 0, st, tables, ntables, nb);
        MPSG_CHECK(hipGetLastError());
    }
    return 0;
}

int launch_seg_range(const ModelDev& m, const void* recs, int64_t elsize, int64_t key_off, int key_bytes,
                     uint64_t key_xmask, const int64_t* begin, const int64_t* count, int nseg, float* q0, float* q1,
                     int* nedge, hipStream_t st) {
    if (nseg <= 0) return 0;
    hipLaunchKernelGGL(seg_range_kernel, dim3(nseg), dim3(256), 0, st, m, (const uint8_t*)recs, elsize, key_off,
                       key_bytes, key_xmask, begin, count, q0, q1, nedge);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

int launch_model_sample(const void* recs, int64_t elsize, int64_t key_off, int key_bytes, uint64_t key_xmask,
                        int64_t n, int64_t S, uint64_t seed, const ModelDev& m, uint64_t* skeys, float* x,
                        float* y, hipStream_t st) {
    hipLaunchKernelGGL(model_sample_kernel, dim3(nblk(S, 256)), dim3(256), 0, st, (const uint8_t*)recs, elsize,
                       key_off, key_bytes, key_xmask, n, S, seed, m, skeys, x, y);
    MPSG_CHECK(hipGetLastError());
    return 0;
}

}  // namespace mpsg
