/*
 * ref_glue.c -- thin C glue that drives the *compiled reference* MP-sort code
 * (radixsort.c, stdlib/msort.c, mpsort-omp.c built from the reference tree by
 * oracle/Makefile into oracle/_ref/).  TEST AND BASELINE INFRASTRUCTURE ONLY.
 *
 * The radix callback below restates binding.pyx:81-121 (u8 / i8 / u4 / i4
 * words, each widened to one little-endian uint64) so the reference is driven
 * with the same key semantics as the Python binding.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>
#include <fcntl.h>
#include <omp.h>
#include "mpsort.h"

typedef struct {
    int64_t offset;
    int32_t width;
    int32_t is_signed;
    int32_t nwords;
    int32_t pad;
} ref_key_desc;

static void radix_desc(const void *ptr, void *radix, void *arg)
{
    const ref_key_desc *k = (const ref_key_desc *)arg;
    const unsigned char *p = (const unsigned char *)ptr + k->offset;
    unsigned char *r = (unsigned char *)radix;
    for (int w = 0; w < k->nwords; w++) {
        uint64_t v;
        if (k->width == 8) {
            memcpy(&v, p + 8 * w, 8);
            if (k->is_signed) v += (uint64_t)9223372036854775808ULL;
        } else if (k->is_signed) {
            int32_t s;
            memcpy(&s, p + 4 * w, 4);
            v = (uint64_t)(int64_t)s + (uint64_t)9223372036854775808ULL;
        } else {
            uint32_t u;
            memcpy(&u, p + 4 * w, 4);
            v = u;
        }
        memcpy(r + 8 * w, &v, 8);
    }
}

/* radix given as raw little-endian bytes at an offset (test-issue7.c:12-17
 * style callers build such radixes by hand) */
typedef struct { int64_t offset; int64_t rsize; } ref_raw_desc;
static void radix_raw(const void *ptr, void *radix, void *arg)
{
    const ref_raw_desc *k = (const ref_raw_desc *)arg;
    memcpy(radix, (const unsigned char *)ptr + k->offset, (size_t)k->rsize);
}

void ref_radix_sort(void *base, size_t n, size_t elsize, const ref_key_desc *k)
{
    radix_sort(base, n, elsize, radix_desc, 8 * (size_t)k->nwords, (void *)k);
}

void ref_radix_sort_raw(void *base, size_t n, size_t elsize, const ref_raw_desc *k)
{
    radix_sort(base, n, elsize, radix_raw, (size_t)k->rsize, (void *)k);
}

/* mpsort_omp prints per-phase timings to stdout (mpsort-omp.c:269-392);
 * they are discarded here. */
void ref_mpsort_omp(void *base, size_t n, size_t elsize, const ref_key_desc *k, int nthreads)
{
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (devnull >= 0) { dup2(devnull, 1); close(devnull); }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    mpsort_omp(base, n, elsize, radix_desc, 8 * (size_t)k->nwords, (void *)k);
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
}

int ref_omp_max_threads(void) { return omp_get_max_threads(); }
