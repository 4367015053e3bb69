// model_dev.h -- device-side evaluation of the learned key-CDF model
// (model.hip trains it): the root MLP's forward pass, F(key) through the
// leaves, and F^-1(q) from the leaves.  Shared by the model kernels and the
// splitter search.
#pragma once
#include <hip/hip_runtime.h>

#include "model.h"

namespace mpsg {

// forward pass with effective weights held in LDS
__device__ __forceinline__ float mlp_forward(const float* __restrict__ e, float x, float* h1, float* h2) {
#pragma unroll
    for (int j = 0; j < kH; j++) h1[j] = tanhf(e[kW1 + j] * x + e[kB1 + j]);
    float r = e[kB3];
#pragma unroll 4
    for (int k = 0; k < kH; k++) {
        float z = e[kB2 + k];
#pragma unroll
        for (int j = 0; j < kH; j++) z += e[kW2 + k * kH + j] * h1[j];
        h2[k] = tanhf(z);
        r += e[kW3 + k] * h2[k];
    }
    return r;
}

__device__ __forceinline__ float mlp_forward_only(const float* __restrict__ e, float x) {
    float h1[kH], h2[kH];
    return mlp_forward(e, x, h1, h2);
}

// ---- evaluation ----------------------------------------------------------
__device__ __forceinline__ float key_to_x(uint64_t key, const ModelDev& m) {
    const uint64_t d = key > m.kmin ? key - m.kmin : 0ull;
    float x = m.log_input ? (float)(log2((double)d + 1.0) * m.log_scale) : (float)((double)d * m.inv_range);
    return x > 1.f ? 1.f : x;
}

__device__ __forceinline__ float model_eval(const float* __restrict__ e, const ModelDev& m, uint64_t key) {
    const float x = key_to_x(key, m);
    const float r = m.linear_root ? x : mlp_forward_only(e, x);
    int l = (int)floorf(r * (float)m.L);
    l = l < 0 ? 0 : (l >= m.L ? m.L - 1 : l);
    const float* lf = m.leaves + 4 * (int64_t)l;
    float f = lf[0] * x + lf[1];
    f = f < lf[2] ? lf[2] : (f > lf[3] ? lf[3] : f);
    return f;
}

// key with x(key) ~ x (the inverse of key_to_x), clamped to [kmin, kmax]
__device__ __forceinline__ uint64_t x_to_key(const ModelDev& m, float x) {
    const double span = (double)(m.kmax - m.kmin);
    double d = m.log_input ? exp2((double)x / m.log_scale) - 1.0 : (double)x * span;
    d = d < 0.0 ? 0.0 : (d > span ? span : d);
    const uint64_t k = m.kmin + (uint64_t)d;
    return k > m.kmax ? m.kmax : k;
}

// F^-1(q) from the leaves alone: the first leaf whose range reaches q, then
// the inverse of its line, kept inside the leaf's x interval.  Monotone in q;
// the boundary tables need no more (bucket membership is an exact key
// comparison whatever the boundaries are).
__device__ inline uint64_t model_inverse_leaves(const ModelDev& m, float q) {
    int lo = 0, hi = m.L;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (m.leaves[4 * (int64_t)mid + 3] < q) lo = mid + 1; else hi = mid;
    }
    if (lo >= m.L) return m.kmax;
    const float* lf = m.leaves + 4 * (int64_t)lo;
    const float x0 = m.xstart[lo], x1 = fminf(m.xstart[lo + 1], 1.0f);
    float x = x0;
    if (lf[2] < q && lf[0] > 0.f) {
        x = (q - lf[1]) / lf[0];
        x = x < x0 ? x0 : (x > x1 ? x1 : x);
    }
    return x_to_key(m, x);
}

}  // namespace mpsg
