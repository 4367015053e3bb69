/*
 * oracle.c -- CPU restatement of MP-sort's histogram-sort hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the MP-sort source tree).  Pinning of each function (how a plausible mistake
 * in it is caught) is listed in tests/test_oracle.py:
 *   oracle_radix         -- edge vectors + numpy closed form
 *   oracle_compar        -- brute force vs Python big-int order
 *   oracle_bisect        -- brute force vs Python floor((a+b)/2)
 *   oracle_radix_sort    -- numpy stable argsort; compiled reference radix_sort
 *   oracle_histogram_sort-- compiled reference mpsort_omp / mpsort_mpi bytes,
 *                           closed forms P_i = k[C[i+1]-1] and brute-force
 *                           per-rank counts, send matrix from the stable order
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define OR_MAXR 64 /* max radix bytes the oracle handles */

typedef struct {
    int64_t offset;   /* byte offset of the key field inside a record */
    int32_t width;    /* bytes per key word: 4 or 8 */
    int32_t is_signed;/* 1 for i4/i8 */
    int32_t nwords;   /* words in the key; the last word is most significant */
    int32_t pad;
} okey_desc;

/* binding.pyx:81-121 -- radix_func_u8/i8/u4/i4: every key word becomes one
 * little-endian uint64; u8 as is, i8 + 2^63, u4 zero-extended, i4 sign-extended
 * then + 2^63.  rsize = 8 * nwords (binding.pyx:206-209). */
void oracle_radix(const void *rec, const okey_desc *k, unsigned char *radix)
{
    const unsigned char *p = (const unsigned char *)rec + k->offset;
    for (int w = 0; w < k->nwords; w++) {
        uint64_t v;
        if (k->width == 8) {
            memcpy(&v, p + 8 * w, 8);
            if (k->is_signed) v += (uint64_t)9223372036854775808ULL;
        } else {
            if (k->is_signed) {
                int32_t s;
                memcpy(&s, p + 4 * w, 4);
                v = (uint64_t)(int64_t)s + (uint64_t)9223372036854775808ULL;
            } else {
                uint32_t u;
                memcpy(&u, p + 4 * w, 4);
                v = u;
            }
        }
        memcpy(radix + 8 * w, &v, 8);
    }
}

/* oracle_radix over n records */
void oracle_radix_batch(const void *base, size_t n, size_t elsize, const okey_desc *k,
                        unsigned char *radix)
{
    size_t rsize = 8 * (size_t)k->nwords;
    for (size_t i = 0; i < n; i++)
        oracle_radix((const unsigned char *)base + i * elsize, k, radix + i * rsize);
}

/* radixsort.c:47-110 + _setup_radix_sort radixsort.c:146-195: for rsize 2/4/8
 * the radix is compared as a native unsigned integer; otherwise word-wise or
 * byte-wise from the highest address down (little-endian host).  All three are
 * the order of the radix bytes read as one little-endian unsigned integer. */
int oracle_compar(const unsigned char *a, const unsigned char *b, size_t rsize)
{
    for (size_t i = rsize; i-- > 0;) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

/* radixsort.c:55-60 (u16/u32/u64: a + ((b-a)>>1)) and radixsort.c:111-144
 * (multi-byte: floor((a+b)/2) with the carry kept as the top bit).  Both are
 * floor((a+b)/2) of little-endian unsigned integers when a <= b. */
void oracle_bisect(unsigned char *mid, const unsigned char *a, const unsigned char *b, size_t rsize)
{
    unsigned char sum[OR_MAXR + 1];
    unsigned carry = 0;
    for (size_t i = 0; i < rsize; i++) {
        unsigned t = (unsigned)a[i] + b[i] + carry;
        sum[i] = (unsigned char)t;
        carry = t >> 8;
    }
    /* shift the (rsize+1)-byte sum right by one bit */
    unsigned hi = carry;
    for (size_t i = rsize; i-- > 0;) {
        unsigned t = (hi << 8) | sum[i];
        mid[i] = (unsigned char)(t >> 1);
        hi = t & 1;
    }
}

/* ---- stable merge sort by radix: stdlib/msort.c:52-174 merges with "<= 0",
 * i.e. on ties the element of the left run goes first (stable).  The oracle
 * sorts an index permutation by precomputed radix and then moves records. */
static void msort_idx(uint32_t *idx, uint32_t *tmp, size_t n,
                      const unsigned char *rad, size_t rsize)
{
    if (n <= 1) return;
    size_t n1 = n / 2, n2 = n - n1;
    uint32_t *b1 = idx, *b2 = idx + n1;
    msort_idx(b1, tmp, n1, rad, rsize);
    msort_idx(b2, tmp, n2, rad, rsize);
    size_t i = 0, j = 0, o = 0;
    while (i < n1 && j < n2) {
        if (oracle_compar(rad + (size_t)b1[i] * rsize, rad + (size_t)b2[j] * rsize, rsize) <= 0)
            tmp[o++] = b1[i++];
        else
            tmp[o++] = b2[j++];
    }
    while (i < n1) tmp[o++] = b1[i++];
    while (j < n2) tmp[o++] = b2[j++];
    memcpy(idx, tmp, n * sizeof(uint32_t));
}

/* radix_sort radixsort.c:35-44: stable sort of nmemb records of elsize bytes by
 * the radix of each record.  In place. */
void oracle_radix_sort(void *base, size_t n, size_t elsize, const okey_desc *k)
{
    if (n <= 1) return;
    size_t rsize = 8 * (size_t)k->nwords;
    unsigned char *rad = malloc(n * rsize);
    uint32_t *idx = malloc(n * sizeof(uint32_t));
    uint32_t *tmp = malloc(n * sizeof(uint32_t));
    unsigned char *copy = malloc(n * elsize);
    for (size_t i = 0; i < n; i++) {
        oracle_radix((unsigned char *)base + i * elsize, k, rad + i * rsize);
        idx[i] = (uint32_t)i;
    }
    msort_idx(idx, tmp, n, rad, rsize);
    memcpy(copy, base, n * elsize);
    for (size_t i = 0; i < n; i++)
        memcpy((unsigned char *)base + i * elsize, copy + (size_t)idx[i] * elsize, elsize);
    free(rad); free(idx); free(tmp); free(copy);
}

/* Same, on precomputed radix bytes (radix callback API, radixsort.c:35-44). */
void oracle_radix_sort_raw(void *base, size_t n, size_t elsize,
                           const unsigned char *radix, size_t rsize)
{
    if (n <= 1) return;
    uint32_t *idx = malloc(n * sizeof(uint32_t));
    uint32_t *tmp = malloc(n * sizeof(uint32_t));
    unsigned char *copy = malloc(n * elsize);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    msort_idx(idx, tmp, n, radix, rsize);
    memcpy(copy, base, n * elsize);
    for (size_t i = 0; i < n; i++)
        memcpy((unsigned char *)base + i * elsize, copy + (size_t)idx[i] * elsize, elsize);
    free(idx); free(tmp); free(copy);
}

/* ---- the histogram sort itself ---------------------------------------- */

typedef struct {
    /* outputs; arrays are caller-allocated with the sizes noted */
    unsigned char *P;        /* (NTask-1) * rsize : final splitters */
    unsigned char *Pmin;     /* rsize */
    unsigned char *Pmax;     /* rsize */
    int64_t *C;              /* NTask+1 */
    int64_t *myCLT;          /* NTask * (NTask+1): per-rank counts of < P */
    int64_t *myCLE;          /* NTask * (NTask+1): per-rank counts of <= P */
    int64_t *myC;            /* NTask * (NTask+1): cumulative send offsets */
    int64_t *send;           /* NTask * NTask: SendCount[sender][receiver] */
    int64_t iterations;      /* bisection rounds */
    int64_t checksum_in;     /* signed-char byte sums */
    int64_t checksum_out;
} oracle_result;

/* internal-parallel.h:8-42 -- _bsearch_last_lt: index of the last item < P in
 * a sorted run, -1 if none. */
static ptrdiff_t bsearch_last_lt(const unsigned char *P, const unsigned char *rad,
                                 size_t n, size_t rsize)
{
    if (n == 0) return -1;
    if (oracle_compar(rad, P, rsize) >= 0) return -1;
    if (oracle_compar(rad + (n - 1) * rsize, P, rsize) < 0) return (ptrdiff_t)n - 1;
    ptrdiff_t l = 0, r = (ptrdiff_t)n - 1;
    while (r > l + 1) {
        ptrdiff_t m = ((r - l + 1) >> 1) + l;
        if (oracle_compar(rad + (size_t)m * rsize, P, rsize) < 0) l = m; else r = m;
    }
    return l;
}

/* internal-parallel.h:49-83 -- _bsearch_last_le. */
static ptrdiff_t bsearch_last_le(const unsigned char *P, const unsigned char *rad,
                                 size_t n, size_t rsize)
{
    if (n == 0) return -1;
    if (oracle_compar(rad, P, rsize) > 0) return -1;
    if (oracle_compar(rad + (n - 1) * rsize, P, rsize) <= 0) return (ptrdiff_t)n - 1;
    ptrdiff_t l = 0, r = (ptrdiff_t)n - 1;
    while (r > l + 1) {
        ptrdiff_t m = ((r - l + 1) >> 1) + l;
        if (oracle_compar(rad + (size_t)m * rsize, P, rsize) <= 0) l = m; else r = m;
    }
    return l;
}

/* internal-parallel.h:97-126 -- _histogram.  Each search starts at the count
 * found for the previous splitter (the reference assumes non-decreasing P). */
static void histogram(const unsigned char *P, int Plength, const unsigned char *rad,
                      size_t n, size_t rsize, int64_t *CLT, int64_t *CLE)
{
    CLT[0] = 0;
    for (int it = 0; it < Plength; it++) {
        int64_t off = CLT[it];
        CLT[it + 1] = bsearch_last_lt(P + (size_t)it * rsize, rad + (size_t)off * rsize,
                                      n - (size_t)off, rsize) + 1 + off;
    }
    CLT[Plength + 1] = (int64_t)n;
    CLE[0] = 0;
    for (int it = 0; it < Plength; it++) {
        int64_t off = CLE[it];
        CLE[it + 1] = bsearch_last_le(P + (size_t)it * rsize, rad + (size_t)off * rsize,
                                      n - (size_t)off, rsize) + 1 + off;
    }
    CLE[Plength + 1] = (int64_t)n;
}

static int64_t checksum_bytes(const void *base, size_t nbytes)
{
    /* mpsort-mpi.c:149-159: sum of the bytes as (signed) char */
    const signed char *p = (const signed char *)base;
    int64_t s = 0;
    for (size_t i = 0; i < nbytes; i++) s += p[i];
    return s;
}

/*
 * mpsort_mpi_newarray_impl + mpsort_mpi_histogram_sort (mpsort-mpi.c:161-604)
 * restated for NTask ranks held in one process.  Rank t owns nin[t] records
 * (consecutive in `base`) and receives nout[t] records (consecutive in `out`).
 * The gather-sort segmenter (mpsort-mpi.c:234-304) is order-preserving and is
 * not restated.
 *
 * `radix` optionally supplies precomputed radix bytes per input record (the
 * radix-callback form); when NULL the key descriptor is used.
 *
 * Returns 0, or -1 for a total size mismatch (mpsort-mpi.c:225-232), -2 for a
 * layout failure (mpsort-mpi.c:707-716), -3 for a checksum mismatch
 * (mpsort-mpi.c:324-330).
 */
int oracle_histogram_sort(const void *base, const int64_t *nin, int NTask,
                          void *out, const int64_t *nout, size_t elsize,
                          const okey_desc *k, const unsigned char *radix_in,
                          size_t rsize_in, oracle_result *res)
{
    size_t rsize = radix_in ? rsize_in : 8 * (size_t)k->nwords;
    int64_t ntot = 0, ntotout = 0;
    for (int t = 0; t < NTask; t++) { ntot += nin[t]; ntotout += nout[t]; }
    if (ntot != ntotout) return -1;
    int NT1 = NTask + 1;

    res->checksum_in = checksum_bytes(base, (size_t)ntot * elsize);

    /* working copies: records and radix of every rank, locally sorted */
    unsigned char *rec = malloc((size_t)ntot * elsize + 1);
    unsigned char *rad = malloc((size_t)ntot * rsize + 1);
    memcpy(rec, base, (size_t)ntot * elsize);
    if (radix_in) memcpy(rad, radix_in, (size_t)ntot * rsize);
    else for (int64_t i = 0; i < ntot; i++) oracle_radix(rec + i * elsize, k, rad + i * rsize);

    int64_t *start = malloc(sizeof(int64_t) * NT1);
    start[0] = 0;
    for (int t = 0; t < NTask; t++) start[t + 1] = start[t] + nin[t];

    /* FirstSort (mpsort-mpi.c:369): stable local sort of each rank; the radix
     * array is permuted along with the records. */
    {
        uint32_t *idx = malloc(sizeof(uint32_t) * (ntot + 1));
        uint32_t *tmp = malloc(sizeof(uint32_t) * (ntot + 1));
        unsigned char *rc = malloc((size_t)ntot * elsize + 1);
        unsigned char *rd = malloc((size_t)ntot * rsize + 1);
        for (int t = 0; t < NTask; t++) {
            int64_t n = nin[t];
            unsigned char *r0 = rad + start[t] * rsize;
            for (int64_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
            msort_idx(idx, tmp, (size_t)n, r0, rsize);
            for (int64_t i = 0; i < n; i++) {
                memcpy(rc + (start[t] + i) * elsize, rec + (start[t] + idx[i]) * elsize, elsize);
                memcpy(rd + (start[t] + i) * rsize, r0 + (size_t)idx[i] * rsize, rsize);
            }
        }
        memcpy(rec, rc, (size_t)ntot * elsize);
        memcpy(rad, rd, (size_t)ntot * rsize);
        free(idx); free(tmp); free(rc); free(rd);
    }

    /* _find_Pmax_Pmin_C (mpsort-mpi.c:606-661) */
    unsigned char *Pmin = res->Pmin, *Pmax = res->Pmax;
    memset(Pmax, 0, rsize);
    memset(Pmin, 0xff, rsize);
    int64_t *C = res->C;
    C[0] = 0;
    for (int t = 0; t < NTask; t++) {
        C[t + 1] = C[t] + nout[t];
        if (nin[t] == 0) continue;
        const unsigned char *first = rad + start[t] * rsize;
        const unsigned char *last = rad + (start[t] + nin[t] - 1) * rsize;
        if (oracle_compar(last, Pmax, rsize) > 0) memcpy(Pmax, last, rsize);
        if (oracle_compar(first, Pmin, rsize) < 0) memcpy(Pmin, first, rsize);
    }
    if (ntot == 0) { memset(Pmin, 0, rsize); memset(Pmax, 0, rsize); }

    /* bisection: piter_* internal-parallel.h:128-248, loop mpsort-mpi.c:381-441 */
    int Pl = NTask - 1;
    unsigned char *P = res->P;
    if (Pl > 0) memset(P, 0, (size_t)Pl * rsize);
    unsigned char *Pleft = malloc((size_t)(Pl + 1) * rsize);
    unsigned char *Pright = malloc((size_t)(Pl + 1) * rsize);
    int *stable = calloc((size_t)Pl + 1, sizeof(int));
    int *narrow = calloc((size_t)Pl + 1, sizeof(int));
    for (int i = 0; i < Pl; i++) {
        memcpy(Pleft + (size_t)i * rsize, Pmin, rsize);
        memcpy(Pright + (size_t)i * rsize, Pmax, rsize);
    }
    int64_t *CLT = calloc(NT1, sizeof(int64_t)), *CLE = calloc(NT1, sizeof(int64_t));
    int64_t *mCLT = malloc(sizeof(int64_t) * NT1), *mCLE = malloc(sizeof(int64_t) * NT1);
    int64_t iter = 0;
    int done = 0;
    while (!done) {
        iter++;
        /* piter_bisect internal-parallel.h:168-198 */
        for (int i = 0; i < Pl; i++) {
            unsigned char *Pi = P + (size_t)i * rsize;
            if (stable[i]) continue;
            if (narrow[i]) {
                memcpy(Pi, Pright + (size_t)i * rsize, rsize);
                stable[i] = 1;
            } else {
                oracle_bisect(Pi, Pleft + (size_t)i * rsize, Pright + (size_t)i * rsize, rsize);
                if (oracle_compar(Pi, Pleft + (size_t)i * rsize, rsize) <= 0) narrow[i] = 1;
            }
        }
        /* _histogram per rank + Allreduce SUM (mpsort-mpi.c:389-394) */
        memset(CLT, 0, sizeof(int64_t) * NT1);
        memset(CLE, 0, sizeof(int64_t) * NT1);
        for (int t = 0; t < NTask; t++) {
            histogram(P, Pl, rad + start[t] * rsize, (size_t)nin[t], rsize, mCLT, mCLE);
            for (int i = 0; i < NT1; i++) { CLT[i] += mCLT[i]; CLE[i] += mCLE[i]; }
        }
        /* piter_accept internal-parallel.h:224-248 */
        for (int i = 0; i < Pl; i++) {
            if (CLT[i + 1] < C[i + 1] && C[i + 1] <= CLE[i + 1]) {
                stable[i] = 1;
                continue;
            }
            if (CLT[i + 1] >= C[i + 1])
                memcpy(Pright + (size_t)i * rsize, P + (size_t)i * rsize, rsize);
            else
                memcpy(Pleft + (size_t)i * rsize, P + (size_t)i * rsize, rsize);
        }
        /* piter_all_done internal-parallel.h:199-216 */
        done = 1;
        for (int i = 0; i < Pl; i++) if (!stable[i]) { done = 0; break; }
    }
    res->iterations = iter;

    /* final _histogram (mpsort-mpi.c:441) */
    for (int t = 0; t < NTask; t++)
        histogram(P, Pl, rad + start[t] * rsize, (size_t)nin[t], rsize,
                  res->myCLT + (size_t)t * NT1, res->myCLE + (size_t)t * NT1);

    /* _solve_for_layout_mpi (mpsort-mpi.c:663-727), one receiver r at a time;
     * myT_CLT[j] = myCLT_j[r+1] after the Alltoall transpose (:450-456). */
    int rc = 0;
    int64_t *myC = res->myC;
    for (int j = 0; j < NTask; j++) myC[(size_t)j * NT1] = 0;
    for (int r = 0; r < NTask && rc == 0; r++) {
        int64_t sure = 0;
        for (int j = 0; j < NTask; j++) {
            myC[(size_t)j * NT1 + r + 1] = res->myCLT[(size_t)j * NT1 + r + 1];
            sure += myC[(size_t)j * NT1 + r + 1];
        }
        int64_t deficit = C[r + 1] - sure;
        for (int j = 0; j < NTask; j++) {
            if (deficit == 0) break;
            if (deficit < 0) { rc = -2; break; }
            int64_t supply = res->myCLE[(size_t)j * NT1 + r + 1] - myC[(size_t)j * NT1 + r + 1];
            if (supply < 0) { rc = -2; break; }
            if (supply <= deficit) { myC[(size_t)j * NT1 + r + 1] += supply; deficit -= supply; }
            else { myC[(size_t)j * NT1 + r + 1] += deficit; deficit = 0; }
        }
    }

    /* SendCount / RecvDispl (mpsort-mpi.c:483-510) and the Alltoallv
     * (mp-mpiu.c:69-156): receiver r gets the runs of senders 0..NTask-1 in
     * sender order. */
    unsigned char *recvbuf = malloc((size_t)ntot * elsize + 1);
    if (rc == 0) {
        int64_t o = 0;
        for (int r = 0; r < NTask; r++) {
            for (int j = 0; j < NTask; j++) {
                int64_t a = myC[(size_t)j * NT1 + r], b = myC[(size_t)j * NT1 + r + 1];
                res->send[(size_t)j * NTask + r] = b - a;
                if (b < a) { rc = -2; break; }
                memcpy(recvbuf + o * elsize, rec + (start[j] + a) * elsize, (size_t)(b - a) * elsize);
                o += b - a;
            }
            if (rc) break;
            if (o != C[r + 1]) { rc = -2; break; }
        }
    }

    /* SecondSort (mpsort-mpi.c:597): stable sort of what each receiver got */
    if (rc == 0) {
        unsigned char *rr = malloc((size_t)ntot * rsize + 1);
        if (radix_in) {
            /* precomputed radix: apply the same exchange to the radix array */
            int64_t o = 0;
            for (int r = 0; r < NTask; r++)
                for (int j = 0; j < NTask; j++) {
                    int64_t a = myC[(size_t)j * NT1 + r], b = myC[(size_t)j * NT1 + r + 1];
                    memcpy(rr + o * rsize, rad + (start[j] + a) * rsize, (size_t)(b - a) * rsize);
                    o += b - a;
                }
        } else {
            for (int64_t i = 0; i < ntot; i++) oracle_radix(recvbuf + i * elsize, k, rr + i * rsize);
        }
        int64_t o = 0;
        uint32_t *idx = malloc(sizeof(uint32_t) * (ntot + 1));
        uint32_t *tmp = malloc(sizeof(uint32_t) * (ntot + 1));
        for (int r = 0; r < NTask; r++) {
            int64_t n = nout[r];
            for (int64_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
            msort_idx(idx, tmp, (size_t)n, rr + o * rsize, rsize);
            for (int64_t i = 0; i < n; i++)
                memcpy((unsigned char *)out + (o + i) * elsize, recvbuf + (o + idx[i]) * elsize, elsize);
            o += n;
        }
        free(idx); free(tmp); free(rr);
        res->checksum_out = checksum_bytes(out, (size_t)ntot * elsize);
        if (res->checksum_out != res->checksum_in) rc = -3;
    }

    free(recvbuf); free(rec); free(rad); free(start);
    free(Pleft); free(Pright); free(stable); free(narrow);
    free(CLT); free(CLE); free(mCLT); free(mCLE);
    return rc;
}

/* ---- the learned path's root network (no reference counterpart) ---------
 * Double-precision restatement of the monotone MLP 1 -> 32 -> 32 -> 1 the GPU
 * trains: w1, W2, w3 = softplus(raw) >= 0, tanh activations, loss = mean of
 * (r(x_i) - y_i)^2.  Parameter layout: w1[32] b1[32] W2[32][32] b2[32] w3[32]
 * b3.  Pinned by central finite differences (tests/test_oracle.py). */
#include <math.h>
#define OM_H 32
#define OM_W1 0
#define OM_B1 (OM_W1 + OM_H)
#define OM_W2 (OM_B1 + OM_H)
#define OM_B2 (OM_W2 + OM_H * OM_H)
#define OM_W3 (OM_B2 + OM_H)
#define OM_B3 (OM_W3 + OM_H)
#define OM_N (OM_B3 + 1)

static double om_softplus(double v) { return v > 30 ? v : log1p(exp(v)); }
static double om_sigmoid(double v) { return 1.0 / (1.0 + exp(-v)); }

double oracle_mlp_forward(const double *raw, double x)
{
    double h1[OM_H], r = raw[OM_B3];
    for (int j = 0; j < OM_H; j++) h1[j] = tanh(om_softplus(raw[OM_W1 + j]) * x + raw[OM_B1 + j]);
    for (int k = 0; k < OM_H; k++) {
        double z = raw[OM_B2 + k];
        for (int j = 0; j < OM_H; j++) z += om_softplus(raw[OM_W2 + k * OM_H + j]) * h1[j];
        r += om_softplus(raw[OM_W3 + k]) * tanh(z);
    }
    return r;
}

/* loss and its gradient with respect to the raw parameters */
double oracle_mlp_grad(const double *raw, const float *x, const float *y, int64_t n, double *grad)
{
    double w1[OM_H], W2[OM_H * OM_H], w3[OM_H];
    for (int j = 0; j < OM_H; j++) { w1[j] = om_softplus(raw[OM_W1 + j]); w3[j] = om_softplus(raw[OM_W3 + j]); }
    for (int q = 0; q < OM_H * OM_H; q++) W2[q] = om_softplus(raw[OM_W2 + q]);
    for (int q = 0; q < OM_N; q++) grad[q] = 0;
    double loss = 0;
    for (int64_t i = 0; i < n; i++) {
        double h1[OM_H], h2[OM_H], r = raw[OM_B3];
        for (int j = 0; j < OM_H; j++) h1[j] = tanh(w1[j] * x[i] + raw[OM_B1 + j]);
        for (int k = 0; k < OM_H; k++) {
            double z = raw[OM_B2 + k];
            for (int j = 0; j < OM_H; j++) z += W2[k * OM_H + j] * h1[j];
            h2[k] = tanh(z);
            r += w3[k] * h2[k];
        }
        double e = r - y[i];
        loss += e * e / n;
        double dy = 2 * e / n, dz1[OM_H];
        for (int j = 0; j < OM_H; j++) dz1[j] = 0;
        grad[OM_B3] += dy;
        for (int k = 0; k < OM_H; k++) {
            grad[OM_W3 + k] += dy * h2[k] * om_sigmoid(raw[OM_W3 + k]);
            double dz2 = dy * w3[k] * (1 - h2[k] * h2[k]);
            grad[OM_B2 + k] += dz2;
            for (int j = 0; j < OM_H; j++) {
                grad[OM_W2 + k * OM_H + j] += dz2 * h1[j] * om_sigmoid(raw[OM_W2 + k * OM_H + j]);
                dz1[j] += dz2 * W2[k * OM_H + j];
            }
        }
        for (int j = 0; j < OM_H; j++) {
            double d1 = dz1[j] * (1 - h1[j] * h1[j]);
            grad[OM_B1 + j] += d1;
            grad[OM_W1 + j] += d1 * x[i] * om_sigmoid(raw[OM_W1 + j]);
        }
    }
    return loss;
}

double oracle_mlp_loss(const double *raw, const float *x, const float *y, int64_t n)
{
    double loss = 0;
    for (int64_t i = 0; i < n; i++) {
        double e = oracle_mlp_forward(raw, x[i]) - y[i];
        loss += e * e / n;
    }
    return loss;
}
