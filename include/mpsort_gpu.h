/*
 * mpsort_gpu.h -- C ABI of the MI355X histogram-sort library (libmpsort_gpu.so).
 *
 * Plain C types only: pointers, sizes, int64 counts.  One process drives one
 * GPU; the P ranks of the reference's MPI sort are P "virtual ranks": P
 * consecutive segments of one input array and P consecutive segments of one
 * output array.  The output is the reference's output: the records of all
 * virtual ranks, stably sorted by radix, cut into the requested per-rank
 * output counts.
 *
 * Reference interfaces replaced (paths relative to the MP-sort tree):
 *   mpsort_gpu_newarray     mpsort_mpi_newarray_impl  mpsort.h:36-47,
 *                           mpsort-mpi.c:161-331 (with the radix callback
 *                           replaced by the key descriptor that
 *                           radix_data_init builds, binding.pyx:50-79)
 *   mpsort_gpu_newarray_radix  the same call with an opaque radix callback
 *   mpsort_gpu_last_stats / mpsort_gpu_report_last_run
 *                           mpsort_mpi_report_last_run mpsort.h:49,
 *                           mpsort-mpi.c:110-119 (phase timers)
 *   mpsort_gpu_last_counts  no exported counterpart: the splitters, C, CLT/CLE
 *                           and send counts the reference keeps on the stack
 *                           of mpsort_mpi_histogram_sort (mpsort-mpi.c:333-364)
 *   mpsort_gpu_set_allocator  MPIU_SetMalloc mp-mpiu.h:4-16 (device side)
 *
 * All entry points return MPSORT_OK (0) or a negative MPSORT_ERR_* code; the
 * reference aborts instead (mpsort-mpi.c:74,185,231,329,500,509,707-716).
 * mpsort_gpu_last_error() gives the message of the last failure.
 * The library is not reentrant (neither is the reference: mpsort-mpi.c:17,105).
 */
#ifndef MPSORT_GPU_H
#define MPSORT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPSORT_OK 0
#define MPSORT_ERR_ARG (-1)          /* invalid argument */
#define MPSORT_ERR_SIZE_MISMATCH (-2)/* sum(nin) != sum(nout): mpsort-mpi.c:225-232 */
#define MPSORT_ERR_LAYOUT (-3)       /* layout solver inconsistency: mpsort-mpi.c:496-510,707-716 */
#define MPSORT_ERR_CHECKSUM (-4)     /* data changed: mpsort-mpi.c:324-330 */
#define MPSORT_ERR_HIP (-5)          /* HIP runtime failure */
#define MPSORT_ERR_NOMEM (-6)        /* device allocation failed */
#define MPSORT_ERR_UNSUPPORTED (-7)  /* key or record shape not supported */
#define MPSORT_ERR_NODEVICE (-8)     /* no GPU */
#define MPSORT_ERR_TIMEOUT (-9)      /* a device wait exceeded its spin limit */

/* Key of a record: nwords words of width 4 or 8 bytes starting at byte
 * `offset`; the last word is the most significant (binding.pyx:50-121,
 * lexsort convention).  Signed words order as signed integers. */
typedef struct {
    int64_t offset;
    int32_t width;     /* 4 or 8 */
    int32_t is_signed; /* 0 or 1 */
    int32_t nwords;    /* >= 1; width*nwords <= 28 unless the record itself
                          is 4/8/16/32 bytes and holds only the key at a
                          4-byte aligned offset */
    int32_t reserved;
} mpsort_key_desc;

/* flags */
#define MPSORT_GPU_DEVICE_PTRS (1u << 0) /* base and out are device pointers */
#define MPSORT_GPU_NO_CHECKSUM (1u << 1) /* skip the byte-sum invariant check */
#define MPSORT_GPU_TIMING (1u << 2)      /* record per-pass kernel times */
#define MPSORT_GPU_LEARNED (1u << 3)     /* learned-CDF splitters + bucket sort */
#define MPSORT_GPU_FORCE_GENERAL (1u << 4) /* always sort packed {key, index} records */
#define MPSORT_GPU_LEARNED_EXCHANGE (1u << 5) /* explicit exchange, then the SecondSort
                                                 (default: with at most 64 ranks and keys of at
                                                 most 8 bytes, each receiver bucket is gathered
                                                 from the senders' sorted runs); either path */
#define MPSORT_GPU_EXCHANGE MPSORT_GPU_LEARNED_EXCHANGE
/* The exchange's run table (MPIU_Alltoallv mp-mpiu.c:125-136): dense walks
 * every (sender, receiver) slot, sparse only the non-empty runs.  Neither
 * flag: the reference's AUTO vote (dense when some sender has more than 128
 * non-empty receivers), unless mpsort_mpi_set_options holds
 * MPSORT_DISABLE_SPARSE_ALLTOALLV or MPSORT_REQUIRE_SPARSE_ALLTOALLV (which
 * wins, as in mpsort-mpi.c:576-583).  Sparse falls back to dense when the
 * non-empty runs exceed 64 per rank on average.  A table chosen by either
 * flag or by a recorded option implies MPSORT_GPU_EXCHANGE.  The result is
 * the same. */
#define MPSORT_GPU_EXCHANGE_DENSE (1u << 6)
#define MPSORT_GPU_EXCHANGE_SPARSE (1u << 7)
/* The plain path's FirstSort (mpsort-mpi.c:369, a stable local sort of
 * every rank): by default, ranks of at least 2^18 records (and at most
 * 4.7 million: 1024 buckets of the in-LDS capacity's half) with keys of at
 * most 8 bytes in 8- or 16-byte sort records, at most 256 ranks, are
 * bucket-sorted -- 1024 buckets per rank from a sorted sample of its keys,
 * a scatter, an in-LDS sort of every bucket; otherwise LSD radix passes.
 * FIRST_RADIX keeps the radix passes; FIRST_BUCKETS takes the bucket sort
 * below the size threshold too.  The result is the same. */
#define MPSORT_GPU_FIRST_RADIX (1u << 8)
#define MPSORT_GPU_FIRST_BUCKETS (1u << 9)

/*
 * Sort the P virtual ranks of `base` (rank t holds nin[t] records, in rank
 * order) into `out` (rank t receives nout[t] records).  elsize is the record
 * size in bytes.  out may equal base (in-place, as mpsort_mpi_impl).  With
 * MPSORT_GPU_DEVICE_PTRS the pointers are device memory; otherwise host
 * memory that is staged through the device.
 */
int mpsort_gpu_newarray(const void *base, const int64_t *nin, int P,
                        void *out, const int64_t *nout, size_t elsize,
                        const mpsort_key_desc *key, unsigned flags);

/* Size of the array behind mpsort_gpu_newarray_noalloc (each GPU the library
 * runs on holds one until the process ends). */
#define MPSORT_GPU_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code between HIP graph captures, tight serving loops, harnesses
 * that forbid hipMalloc: mpsort_gpu_newarray on DEVICE pointers (the flag is
 * implied), plain path, with every device workspace carved from a fixed
 * array of MPSORT_GPU_NOALLOC_SCRATCH_BYTES per GPU instead of allocated.  It
 * calls no hipMalloc, hipFree or other allocating routine.  A sort whose
 * workspace does not fit (about 2 x n sort records plus tables: roughly 4
 * million 16-byte records over 64 ranks) fails with MPSORT_ERR_NOMEM and
 * names its need.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU race.  Output, counts and statistics equal
 * mpsort_gpu_newarray's.
 */
int mpsort_gpu_newarray_noalloc(const void *base, const int64_t *nin, int P,
                                void *out, const int64_t *nout, size_t elsize,
                                const mpsort_key_desc *key, unsigned flags);

/* The same with the reference's opaque radix callback (mpsort.h:1-4): the
 * radix of every record is materialised once on the host (compared as a
 * little-endian unsigned integer, radixsort.c:146-195), the {radix, index}
 * pairs are sorted on the GPU and the records gathered.  Any rsize: up to 28
 * bytes (24 past 2^32 records) the pair is one sort record; a longer radix
 * is sorted by stable passes over its 24-byte chunks, least significant
 * first, and the counts recorded are those of the top chunk's pass.  Host
 * pointers only. */
int mpsort_gpu_newarray_radix(const void *base, const int64_t *nin, int P,
                              void *out, const int64_t *nout, size_t elsize,
                              void (*radix)(const void *ptr, void *radix, void *arg),
                              size_t rsize, void *arg, unsigned flags);

typedef struct {
    double t_first_sort_ms;  /* FirstSort (mpsort-mpi.c:369) */
    double t_findp_ms;       /* PmaxPmin + bisection (mpsort-mpi.c:375-441) */
    double t_laysolve_ms;    /* LayDistr + LaySolve (mpsort-mpi.c:443-510) */
    double t_exchange_ms;    /* Exchange (mpsort-mpi.c:571-592) */
    double t_second_sort_ms; /* SecondSort (mpsort-mpi.c:597) */
    double t_total_ms;       /* device time of the whole call */
    double t_h2d_ms;         /* host staging (host-pointer calls only) */
    double t_d2h_ms;
    int32_t nvranks;
    int32_t record_bytes;    /* sort record size actually moved */
    int32_t first_passes;    /* digit passes run by FirstSort */
    int32_t second_passes;
    int32_t bisect_rounds;   /* max over splitters */
    int32_t general_path;    /* 1 when packed {key, index} records were sorted */
    int64_t nrecords;
    int64_t checksum_in;
    int64_t checksum_out;
    int32_t npass_ms;        /* entries used in pass_ms */
    int32_t learned;         /* 1 when the learned path ran */
    double pass_ms[64];      /* per digit pass kernel time (FirstSort then SecondSort) */
    /* learned path */
    int32_t splitters_exact;   /* proposals accepted without repair */
    int32_t splitters_repaired;
    int32_t buckets;
    int32_t buckets_overflow;  /* buckets that fell back to the plain sort */
    int32_t repair_rounds;
    int32_t plain_sides;       /* local sorts that took the plain radix path */
    double t_model_ms;
    double t_bucket_ms;
    int32_t buckets_lds_radix; /* buckets sorted by the in-LDS radix sort */
    int32_t buckets_overflow_first; /* of buckets_overflow, those of the FirstSort */
    int32_t buckets_listed;    /* buckets sorted by the 1024-thread in-LDS form from a list */
    int32_t exchange_sparse;   /* 1 when the exchange walked the non-empty runs only */
    int32_t exchange_runs;     /* runs in the table it walked (P * P when dense) */
    int32_t second_merged;     /* 1 when the SecondSort gathered every receiver bucket from the
                                  senders' sorted segments (no exchange, no digit passes) */
} mpsort_gpu_stats;

int mpsort_gpu_last_stats(mpsort_gpu_stats *stats);

/* Fault hooks, for tests of the paths that the checks of the reference
 * guard (it aborts instead, mpsort-mpi.c:324-330,496-510,707-716); no
 * counterpart in the reference.  They stay set for the following calls of
 * this process until cleared with faults = 0.
 *   MPSORT_GPU_FAULT_SPLITTERS   learned path: the model's first proposal for
 *       every splitter is moved by 0.37 in quantile, so the repair rounds
 *       find every splitter
 *   MPSORT_GPU_FAULT_BUCKET_CAP  learned path: buckets above bucket_cap
 *       records (>= 1) take the overflow fallback, as if above the in-LDS
 *       capacity
 *   MPSORT_GPU_FAULT_PLAIN       learned path: every local sort takes the
 *       plain radix sort of its whole side
 *   MPSORT_GPU_FAULT_CORRUPT     mpsort_gpu_newarray / _learned: a byte past
 *       the key of one record is flipped between FirstSort and SecondSort
 *       (plain path, more than one rank, key not filling the record), else
 *       the recorded input sum is changed; the call then fails with
 *       MPSORT_ERR_CHECKSUM (unless MPSORT_GPU_NO_CHECKSUM) */
#define MPSORT_GPU_FAULT_SPLITTERS (1u << 0)
#define MPSORT_GPU_FAULT_BUCKET_CAP (1u << 1)
#define MPSORT_GPU_FAULT_PLAIN (1u << 2)
#define MPSORT_GPU_FAULT_CORRUPT (1u << 3)
int mpsort_gpu_set_faults(unsigned faults, int64_t bucket_cap);

/* Prints the phase times of the last call in the reference's format
 * (mpsort-mpi.c:110-119). */
void mpsort_gpu_report_last_run(void);

/*
 * Per-virtual-rank counts of the last mpsort_gpu_newarray call, in the
 * reference's layout (P = nvranks):
 *   splitters  (P-1) radix values of rsize bytes each (rsize = 8*nwords for
 *              the descriptor form, the callback's rsize otherwise)
 *   pmin/pmax  rsize bytes each
 *   C          P+1 cumulative output counts
 *   myCLT/myCLE  P x (P+1): rank j's counts of radix < / <= splitter i at
 *              [j*(P+1) + i+1], with columns 0 and P fixed at 0 and n_j
 *   myC        P x (P+1): cumulative send offsets of rank j
 *   send       P x P: records rank j sends to rank r at [j*P + r]
 * Any pointer may be NULL.
 */
int mpsort_gpu_last_counts(unsigned char *splitters, unsigned char *pmin, unsigned char *pmax,
                           int64_t *C, int64_t *myCLT, int64_t *myCLE, int64_t *myC,
                           int64_t *send, int64_t *rsize_out);

/* ---- learned-splitter path ------------------------------------------------
 * No reference counterpart: a learned key-CDF model proposes the P-1
 * splitters (checked with exact counts and repaired by the reference's
 * bisection inside the proposed bracket) and the bucket boundaries of a
 * one-pass bucket scatter + in-LDS bucket sort that replaces the radix passes
 * of both local sorts.  The output is bit-identical to mpsort_gpu_newarray.
 *
 * The model is trained on nsample records sampled from `sorted`, the output of
 * one plain sort (their exact ranks are their positions), on the GPU.  Keys of
 * at most 8 bytes. */
typedef struct mpsort_gpu_model_s mpsort_gpu_model;
#define MPSORT_GPU_MODEL_LINEAR_ROOT (1u << 8) /* no network: linear leaf routing */

typedef struct {
    int64_t nsample;
    int32_t leaves;
    int32_t linear_root;
    int32_t steps;           /* Adam steps run */
    int32_t nheavy;          /* duplicate-heavy keys kept as exact boundaries */
    int32_t max_leaf_samples;/* largest number of samples routed to one leaf */
    int32_t log_input;       /* 1: root input log2(1 + key - kmin) scaled to [0, 1] */
    double train_ms;
    float loss;              /* final mean squared error of the root */
    float err_max;           /* max |F(key) - rank/n| over the samples */
    uint64_t kmin, kmax;
} mpsort_gpu_model_info;

int mpsort_gpu_model_train(mpsort_gpu_model **model, const void *sorted, int64_t n, size_t elsize,
                           const mpsort_key_desc *key, int64_t nsample, uint64_t seed, unsigned flags);
int mpsort_gpu_model_free(mpsort_gpu_model *model);
int mpsort_gpu_model_get_info(const mpsort_gpu_model *model, mpsort_gpu_model_info *info);
/* F(key) for n keys given as the order-preserving uint64 (host arrays) */
int mpsort_gpu_model_eval(const mpsort_gpu_model *model, const uint64_t *keys, int64_t n, float *out);
/* raw MLP parameters (1153 floats: w1[32] b1[32] W2[32][32] b2[32] w3[32] b3)
 * and the leaves (4 floats each: slope, intercept, lo, hi); either may be NULL */
int mpsort_gpu_model_params(const mpsort_gpu_model *model, float *raw, float *leaves);
/* gradient of the root's mean squared error at raw parameters `raw` over n
 * samples (x, y) computed by the training kernels (host arrays) */
int mpsort_gpu_mlp_grad(const float *raw, const float *x, const float *y, int64_t n, float *grad,
                        float *loss);
int mpsort_gpu_newarray_learned(const void *base, const int64_t *nin, int P, void *out,
                                const int64_t *nout, size_t elsize, const mpsort_key_desc *key,
                                const mpsort_gpu_model *model, unsigned flags);

/* Batched segmented sort: many independent stable sorts (the reference's
 * counterpart is a loop of radix_sort, radixsort.c:35-44, one call per
 * segment).  A plan holds groups; a group is nseg segments of one record type
 * laid out back to back in device memory (d_in), sorted into d_out (may equal
 * d_in).  Records of 4/8/16/32 bytes, key 4-byte aligned inside the record.
 * MPSORT_GPU_SEGPLAN_GRAPH captures the whole plan once in a hipGraph and
 * replays it on every run (the segment sizes and pointers are fixed by the
 * plan); MPSORT_GPU_SEGPLAN_TUNE times tile shapes and the in-LDS size
 * threshold, then the workgroup shape of the in-LDS segments (the waves that
 * share a CU), on the first run and caches the choice per (record size, key
 * width, size class) in memory and in $MPSORT_GPU_TUNE_CACHE (default
 * ~/.cache/mpsort_gpu_segsort.tune). */
typedef struct mpsort_gpu_segplan_s mpsort_gpu_segplan;
#define MPSORT_GPU_SEGPLAN_GRAPH (1u << 0)
#define MPSORT_GPU_SEGPLAN_TUNE (1u << 1)
/* the workgroup shape of the segments sorted in LDS when not tuned (default:
 * 1024 threads x 9 records, one workgroup per CU): 512 x 9 (two per CU,
 * segments up to 4608 records) or 512 x 6 (three per CU, up to 3072); larger
 * ones below the threshold take the 1024-thread form */
#define MPSORT_GPU_SEGPLAN_SMALL_512x9 (1u << 2)
#define MPSORT_GPU_SEGPLAN_SMALL_512x6 (1u << 3)
/* every segment sorted by launches of its own, one segment after the other
 * on the stream with no host wait between them (the loop of per-segment
 * sorts the batched form replaces, with the same kernels and workspaces;
 * for comparison) */
#define MPSORT_GPU_SEGPLAN_EACH (1u << 4)
int mpsort_gpu_segplan_create(mpsort_gpu_segplan **plan, unsigned flags);
int mpsort_gpu_segplan_add(mpsort_gpu_segplan *plan, const void *d_in, void *d_out, const int64_t *seg_count,
                           int nseg, size_t elsize, const mpsort_key_desc *key);
/* stats (may be NULL): t_total_ms and nrecords are filled */
int mpsort_gpu_segplan_run(mpsort_gpu_segplan *plan, mpsort_gpu_stats *stats);
/* the tuning a group runs with: segments sorted in LDS / by onesweep, tile */
int mpsort_gpu_segplan_info(const mpsort_gpu_segplan *plan, int group, int *nsmall, int *nlarge, int *block,
                            int *items, int *threshold);
/* the workgroup shape of a group's in-LDS segments (0: 1024 x 9, 1: 512 x 9,
 * 2: 512 x 6) and the waves that then share a CU */
int mpsort_gpu_segplan_shape(const mpsort_gpu_segplan *plan, int group, int *form, int *waves_per_cu);
int mpsort_gpu_segplan_free(mpsort_gpu_segplan *plan);
/* one group, no graph: create + add + run + free */
int mpsort_gpu_sort_segments(const void *d_in, void *d_out, const int64_t *seg_count, int nseg, size_t elsize,
                             const mpsort_key_desc *key, unsigned flags);

/* Device allocator hook (MPIU_SetMalloc mp-mpiu.h:4-16).  NULL restores
 * hipMalloc/hipFree.  Frees every cached workspace. */
typedef void *(*mpsort_gpu_malloc_fn)(size_t bytes, void *userdata);
typedef void (*mpsort_gpu_free_fn)(void *ptr, void *userdata);
int mpsort_gpu_set_allocator(mpsort_gpu_malloc_fn m, mpsort_gpu_free_fn f, void *userdata);

/* Frees the cached device workspaces of the current device. */
int mpsort_gpu_release(void);

/* Number of virtual ranks used by the MPI-shaped entry points
 * (mpsort_mpi_impl / mpsort_mpi_newarray_impl / radix_sort); default 1, or
 * the MPSORT_GPU_VRANKS environment variable. */
void mpsort_gpu_set_vranks(int P);
int mpsort_gpu_get_vranks(void);

const char *mpsort_gpu_last_error(void);
const char *mpsort_gpu_version(void);

/* Device-memory helpers for callers without a HIP toolchain (tests, ctypes). */
void *mpsort_gpu_dmalloc(size_t bytes);
int mpsort_gpu_dfree(void *p);
int mpsort_gpu_memcpy_h2d(void *dst, const void *src, size_t bytes);
int mpsort_gpu_memcpy_d2h(void *dst, const void *src, size_t bytes);
int mpsort_gpu_synchronize(void);
/* device-to-device copy by the library's own copy kernel (a plain streaming
 * kernel of known traffic, used to calibrate hardware byte counters) */
int mpsort_gpu_dcopy(void *dst, const void *src, size_t bytes);
int mpsort_gpu_device_count(void);

/* Option bits of the reference (mpsort.h:17-24).  Accepted and recorded: on
 * one GPU there is no Alltoallv to make sparse and no rank to gather to, so
 * they do not change the result (nor does any of them in the reference). */
void mpsort_mpi_set_options(int options);
int mpsort_mpi_has_options(int options);
void mpsort_mpi_unset_options(int options);

#ifdef __cplusplus
}
#endif
#endif
