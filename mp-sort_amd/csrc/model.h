// model.h -- parameter layout of the learned key-CDF model (model.hip).
#pragma once
#include <stdint.h>

#include "internal.h"

namespace mpsg {

constexpr int kH = 32;  // hidden width of the MLP root
// raw parameter vector: w1[32] b1[32] W2[32][32] b2[32] w3[32] b3
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kH;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH;
constexpr int kMlpParams = kB3 + 1;

// device view of a trained model
struct ModelDev {
    const float* eff;     // kMlpParams effective weights
    const float* leaves;  // L x {a, b, lo, hi}
    int L;
    int linear_root;      // 1: leaf = floor(x * L) (no network)
    uint64_t kmin;
    double inv_range;     // 1 / (kmax - kmin)
    int log_input;        // 1: x = log2(1 + key - kmin) * log_scale
    double log_scale;     // 1 / log2(1 + kmax - kmin)
    const float* xstart;  // L + 1: smallest x routed to each leaf (xstart[L] = 2)
    uint64_t kmax;
};

int launch_mlp_effective(const float* raw, float* eff, hipStream_t st);
int mlp_grad_blocks(int64_t n, int64_t stride, int64_t offset);
int launch_mlp_grad(const float* x, const float* y, int64_t n, int64_t stride, int64_t offset,
                    const float* raw, const float* eff, float* partial, float* loss_partial, hipStream_t st);
int launch_mlp_adam(float* raw, const float* partial, int nblocks, float* m, float* v, float lr, int step,
                    float* grad_out, hipStream_t st);
int launch_leaves(const float* x, const float* y, int64_t S, const float* eff, int L, int linear_root,
                  int* leaf, int64_t* start, float* leaves, hipStream_t st);
// xstart[l] = smallest x the root routes to leaf l or beyond
int launch_leaf_xstart(const float* eff, int L, int linear_root, float* xstart, hipStream_t st);
int launch_model_eval(const ModelDev& m, const uint64_t* keys, int64_t n, float* out, hipStream_t st);
// nedge (optional): per table, boundaries kept on each side of [q0, q1];
// heavy keys (ascending, at most kMaxHeavy) with their sampled quantile
// ranges {qa, qb}: a boundary whose quantile lies in such a range is that
// key, and a table isolates the heavy keys inside its span that weigh at
// least one of its buckets; ordered: the tables are of consecutive quantile
// ranges and are made to read as one sorted sequence
constexpr int kMaxHeavy = 16384;
int launch_model_tables(const ModelDev& m, uint64_t kmax, const float* q0, const float* q1, int ntables,
                        int nb, uint64_t* tables, const uint64_t* heavy, const float2* heavy_q, int nheavy,
                        hipStream_t st, const int* nedge = nullptr, bool ordered = false);
// per segment of unsorted records: the quantile range of its keys and its
// edge boundaries (see model_tables_kernel)
int launch_seg_range(const ModelDev& m, const void* recs, int64_t elsize, int64_t key_off, int key_bytes,
                     uint64_t key_xmask, const int64_t* begin, const int64_t* count, int nseg, float* q0, float* q1,
                     int* nedge, hipStream_t st);
int launch_model_sample(const void* recs, int64_t elsize, int64_t key_off, int key_bytes, uint64_t key_xmask,
                        int64_t n, int64_t S, uint64_t seed, const ModelDev& m, uint64_t* skeys, float* x,
                        float* y, hipStream_t st);

}  // namespace mpsg
