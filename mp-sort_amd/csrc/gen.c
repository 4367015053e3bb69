/*
 * gen.c -- seeded synthetic workloads for the histogram-sort path.
 *
 * Inputs are generated once on the host and the same bytes are fed to the GPU
 * path and to the CPU checkers.  Every generator is deterministic for a given
 * seed regardless of the thread count: the parallel ones derive each record's
 * random stream from (seed, record index) with the SplitMix64 finaliser, a
 * bijection on 64-bit words (Steele, Lea, Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014).
 *
 * Workloads (sizes are arguments):
 *   mpsort_gen_pcg64      uniform uint64 keys from PCG-XSL-RR-128/64 (O'Neill
 *                         2014), the generator numpy calls PCG64
 *   mpsort_gen_records16  {u64 key, u64 global index} records
 *   mpsort_gen_clustered  63-bit Morton codes of a clustered 3-D field
 *   mpsort_gen_lognormal_mix, mpsort_gen_zipf, mpsort_gen_nearly_sorted
 *   mpsort_gen_uniform    uniform uint64 keys, counter based
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ---- PCG64 (XSL-RR 128/64) ---------------------------------------------- */
typedef unsigned __int128 u128;
static const u128 PCG_MULT = ((u128)0x2360ED051FC65DA4ULL << 64) | 0x4385DF649FCCF645ULL;

/* Fill keys[0..n) with successive PCG64 outputs; the state is advanced before
 * each output (numpy's PCG64.random_raw convention). */
void mpsort_gen_pcg64(uint64_t *keys, size_t n, uint64_t state_hi, uint64_t state_lo,
                      uint64_t inc_hi, uint64_t inc_lo)
{
    u128 s = ((u128)state_hi << 64) | state_lo;
    u128 inc = (((u128)inc_hi << 64) | inc_lo) | 1;
    for (size_t i = 0; i < n; i++) {
        s = s * PCG_MULT + inc;
        uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
        unsigned rot = (unsigned)(s >> 122);
        keys[i] = (x >> rot) | (x << ((64 - rot) & 63));
    }
}

/* ---- counter-based streams ------------------------------------------------ */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t stream_u64(uint64_t seed, uint64_t stream, uint64_t i, uint64_t k)
{
    return mix64(mix64(seed ^ (stream * 0x9E3779B97F4A7C15ULL)) + i * 0xD1B54A32D192ED03ULL
                 + k * 0x9E3779B97F4A7C15ULL);
}
static inline double stream_unit(uint64_t seed, uint64_t stream, uint64_t i, uint64_t k)
{
    return (double)(stream_u64(seed, stream, i, k) >> 11) * 0x1.0p-53; /* [0,1) */
}
static inline double stream_normal(uint64_t seed, uint64_t stream, uint64_t i, uint64_t k)
{
    /* Box-Muller on two stream words */
    double u1 = 1.0 - stream_unit(seed, stream, i, 2 * k);
    double u2 = stream_unit(seed, stream, i, 2 * k + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void mpsort_gen_uniform(uint64_t *keys, size_t n, uint64_t seed)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) keys[i] = stream_u64(seed, 1, i, 0);
}

/* {u64 key, u64 idx}: keys from `keys`, idx = global index (base + i) */
void mpsort_gen_records16(void *out, const uint64_t *keys, size_t n, uint64_t base)
{
    uint64_t *o = (uint64_t *)out;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) { o[2 * i] = keys[i]; o[2 * i + 1] = base + i; }
}

/* ---- clustered Morton field ---------------------------------------------- */
static inline uint64_t spread3(uint64_t v)
{
    /* put bit b of a 21-bit value at bit 3b */
    v &= 0x1FFFFF;
    v = (v | (v << 32)) & 0x001F00000000FFFFULL;
    v = (v | (v << 16)) & 0x001F0000FF0000FFULL;
    v = (v | (v << 8)) & 0x100F00F00F00F00FULL;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}
uint64_t mpsort_morton63(uint32_t qx, uint32_t qy, uint32_t qz)
{
    return spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
}
static inline uint32_t quant21(double x)
{
    x = x - floor(x); /* periodic wrap into [0,1) */
    double q = floor(x * 2097152.0);
    if (q < 0) q = 0;
    if (q > 2097151.0) q = 2097151.0;
    return (uint32_t)q;
}

/*
 * Morton keys of n points: nclump Gaussian clumps with centres uniform in the
 * unit box and radii median_r * exp(sigma_ln * N(0,1)); a fraction
 * frac_uniform of the points is uniform background; the clump of a clumped
 * point is chosen uniformly.  Positions wrap periodically and are quantised to
 * 21 bits per axis.  Writes {u64 morton, i64 index} records when rec != NULL,
 * and the keys alone when keys != NULL.
 */
/* The field (clump centres and radii) comes from field_seed, the points
 * from seed: two seeds with one field_seed are two realisations of one
 * clustered field. */
void mpsort_gen_clustered_field(uint64_t *keys, void *rec, size_t n, uint64_t field_seed, uint64_t seed,
                                int nclump, double median_r, double sigma_ln, double frac_uniform)
{
    double *cx = malloc(sizeof(double) * 4 * (size_t)nclump);
    for (int c = 0; c < nclump; c++) {
        cx[4 * c + 0] = stream_unit(field_seed, 10, (uint64_t)c, 0);
        cx[4 * c + 1] = stream_unit(field_seed, 10, (uint64_t)c, 1);
        cx[4 * c + 2] = stream_unit(field_seed, 10, (uint64_t)c, 2);
        cx[4 * c + 3] = median_r * exp(sigma_ln * stream_normal(field_seed, 11, (uint64_t)c, 0));
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        double x, y, z;
        if (stream_unit(seed, 12, i, 0) < frac_uniform) {
            x = stream_unit(seed, 13, i, 0);
            y = stream_unit(seed, 13, i, 1);
            z = stream_unit(seed, 13, i, 2);
        } else {
            uint64_t c = (uint64_t)(stream_unit(seed, 14, i, 0) * nclump);
            if (c >= (uint64_t)nclump) c = (uint64_t)nclump - 1;
            double r = cx[4 * c + 3];
            x = cx[4 * c + 0] + r * stream_normal(seed, 15, i, 0);
            y = cx[4 * c + 1] + r * stream_normal(seed, 15, i, 1);
            z = cx[4 * c + 2] + r * stream_normal(seed, 15, i, 2);
        }
        uint64_t m = mpsort_morton63(quant21(x), quant21(y), quant21(z));
        if (keys) keys[i] = m;
        if (rec) { ((uint64_t *)rec)[2 * i] = m; ((uint64_t *)rec)[2 * i + 1] = i; }
    }
    free(cx);
}

/* the field and the points from one seed */
void mpsort_gen_clustered(uint64_t *keys, void *rec, size_t n, uint64_t seed,
                          int nclump, double median_r, double sigma_ln, double frac_uniform)
{
    mpsort_gen_clustered_field(keys, rec, n, seed, seed, nclump, median_r, sigma_ln, frac_uniform);
}

/* A radix callback of the reference's form (mpsort.h:1-4) for tests and
 * benchmarks of the callback path: arg points to {int64 offset, int64
 * nwords}; the radix is the nwords little-endian u64 words at offset. */
void mpsort_radix_u64_words(const void *ptr, void *radix, void *arg)
{
    const int64_t *a = (const int64_t *)arg;
    memcpy(radix, (const char *)ptr + a[0], 8 * (size_t)a[1]);
}

/* log-space mixture of 8 Gaussians cast to uint64 (clipped to the range) */
void mpsort_gen_lognormal_mix(uint64_t *keys, size_t n, uint64_t seed)
{
    static const double mu[8] = {4.0, 9.0, 14.0, 19.0, 24.0, 29.0, 34.0, 39.0};
    static const double sg[8] = {0.3, 0.5, 0.8, 1.0, 0.6, 0.4, 1.2, 0.7};
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        int c = (int)(stream_unit(seed, 20, i, 0) * 8.0);
        if (c > 7) c = 7;
        double v = exp(mu[c] + sg[c] * stream_normal(seed, 21, i, 0));
        uint64_t k;
        if (!(v >= 0.0)) k = 0;
        else if (v >= 18446744073709551615.0) k = UINT64_MAX;
        else k = (uint64_t)v;
        keys[i] = k;
    }
}

/* Zipf(s) ranks over nvalues distinct values, each rank r (1-based) mapped to
 * a key by the SplitMix64 finaliser (a bijection on uint64). */
void mpsort_gen_zipf(uint64_t *keys, size_t n, uint64_t seed, double s, uint32_t nvalues)
{
    double *cdf = malloc(sizeof(double) * nvalues);
    double acc = 0.0;
    for (uint32_t r = 0; r < nvalues; r++) { acc += pow((double)(r + 1), -s); cdf[r] = acc; }
    for (uint32_t r = 0; r < nvalues; r++) cdf[r] /= acc;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        double u = stream_unit(seed, 30, i, 0);
        uint32_t lo = 0, hi = nvalues - 1;
        while (lo < hi) { uint32_t m = (lo + hi) / 2; if (cdf[m] > u) hi = m; else lo = m + 1; }
        keys[i] = mix64((uint64_t)lo + 1);
    }
    free(cdf);
}

/* sorted ramp over the full uint64 range with npairs random position swaps */
void mpsort_gen_nearly_sorted(uint64_t *keys, size_t n, uint64_t seed, size_t npairs)
{
    uint64_t step = n ? (UINT64_MAX / n) : 0;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) keys[i] = (uint64_t)i * step;
    for (size_t p = 0; p < npairs && n > 1; p++) {
        size_t a = (size_t)(stream_u64(seed, 40, p, 0) % n);
        size_t b = (size_t)(stream_u64(seed, 40, p, 1) % n);
        uint64_t t = keys[a]; keys[a] = keys[b]; keys[b] = t;
    }
}

/* Fisher-Yates permutation of 0..n-1 driven by the counter stream */
void mpsort_gen_permutation(uint64_t *out, size_t n, uint64_t seed)
{
    for (size_t i = 0; i < n; i++) out[i] = i;
    for (size_t i = n; i > 1; i--) {
        size_t j = (size_t)(stream_u64(seed, 50, i, 0) % i);
        uint64_t t = out[i - 1]; out[i - 1] = out[j]; out[j] = t;
    }
}

/* segment sizes: log-uniform on [lo, hi], then rescaled so they sum to total
 * (each at least 1); returns the sum actually produced */
uint64_t mpsort_gen_segment_sizes(uint64_t *sizes, size_t nseg, uint64_t seed,
                                  double lo, double hi, uint64_t total)
{
    double *raw = malloc(sizeof(double) * nseg);
    double sum = 0.0;
    for (size_t i = 0; i < nseg; i++) {
        raw[i] = exp(log(lo) + (log(hi) - log(lo)) * stream_unit(seed, 60, i, 0));
        sum += raw[i];
    }
    uint64_t acc = 0;
    for (size_t i = 0; i < nseg; i++) {
        double v = total ? floor(raw[i] * (double)total / sum) : floor(raw[i]);
        sizes[i] = v < 1.0 ? 1 : (uint64_t)v;
        acc += sizes[i];
    }
    free(raw);
    return acc;
}

uint64_t mpsort_gen_mix64(uint64_t z) { return mix64(z); }
