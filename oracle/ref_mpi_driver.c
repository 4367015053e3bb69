/*
 * ref_mpi_driver.c -- runs the compiled reference mpsort_mpi_newarray
 * (mpsort-mpi.c + mp-mpiu.c from the reference tree, built by oracle/Makefile
 * into oracle/_ref/ref_mpi) under mpiexec.  TEST AND BASELINE INFRASTRUCTURE.
 *
 * usage: mpiexec -n P ref_mpi IN OUT ELSIZE OFFSET WIDTH SIGNED NWORDS REPEAT \
 *                              nin_0..nin_{P-1} nout_0..nout_{P-1}
 * IN holds sum(nin) records; rank t reads its slice, the sorted slices are
 * gathered to rank 0 and written to OUT.  Rank 0 prints the median wall time
 * of REPEAT calls (barrier + MPI_Wtime around the call, as bench-mpi.c:205-214).
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include "mpsort.h"

typedef struct { int64_t offset; int32_t width, is_signed, nwords, pad; } key_desc;

static void radix_desc(const void *ptr, void *radix, void *arg)
{
    const key_desc *k = (const key_desc *)arg;
    const unsigned char *p = (const unsigned char *)ptr + k->offset;
    unsigned char *r = (unsigned char *)radix;
    for (int w = 0; w < k->nwords; w++) {
        uint64_t v;
        if (k->width == 8) {
            memcpy(&v, p + 8 * w, 8);
            if (k->is_signed) v += (uint64_t)9223372036854775808ULL;
        } else if (k->is_signed) {
            int32_t s; memcpy(&s, p + 4 * w, 4);
            v = (uint64_t)(int64_t)s + (uint64_t)9223372036854775808ULL;
        } else {
            uint32_t u; memcpy(&u, p + 4 * w, 4); v = u;
        }
        memcpy(r + 8 * w, &v, 8);
    }
}

static int cmpd(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc != 10 + 2 * size) {
        if (rank == 0) fprintf(stderr, "bad arguments for %d ranks\n", size);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    const char *fin = argv[1], *fout = argv[2];
    size_t elsize = (size_t)atoll(argv[3]);
    key_desc k;
    k.offset = atoll(argv[4]); k.width = atoi(argv[5]); k.is_signed = atoi(argv[6]);
    k.nwords = atoi(argv[7]); k.pad = 0;
    int repeat = atoi(argv[8]);
    int quiet = atoi(argv[9]);
    int64_t *nin = malloc(sizeof(int64_t) * size), *nout = malloc(sizeof(int64_t) * size);
    int64_t off = 0, ooff = 0, ntot = 0;
    for (int t = 0; t < size; t++) {
        nin[t] = atoll(argv[10 + t]);
        nout[t] = atoll(argv[10 + size + t]);
        if (t < rank) { off += nin[t]; ooff += nout[t]; }
        ntot += nin[t];
    }
    unsigned char *orig = malloc(nin[rank] * elsize + 1);
    unsigned char *mine = malloc(nin[rank] * elsize + 1);
    unsigned char *out = malloc(nout[rank] * elsize + 1);
    FILE *f = fopen(fin, "rb");
    if (!f) MPI_Abort(MPI_COMM_WORLD, 3);
    fseek(f, (long)(off * elsize), SEEK_SET);
    if (fread(orig, elsize, nin[rank], f) != (size_t)nin[rank]) MPI_Abort(MPI_COMM_WORLD, 4);
    fclose(f);
    if (quiet) mpsort_mpi_set_options(0);

    double *times = malloc(sizeof(double) * (repeat > 0 ? repeat : 1));
    for (int it = 0; it < repeat; it++) {
        memcpy(mine, orig, nin[rank] * elsize);
        MPI_Barrier(MPI_COMM_WORLD);
        double t0 = MPI_Wtime();
        mpsort_mpi_newarray(mine, nin[rank], out, nout[rank], elsize,
                            radix_desc, 8 * (size_t)k.nwords, &k, MPI_COMM_WORLD);
        MPI_Barrier(MPI_COMM_WORLD);
        times[it] = MPI_Wtime() - t0;
    }
    /* gather outputs to rank 0 */
    int *cnt = malloc(sizeof(int) * size), *dsp = malloc(sizeof(int) * size);
    for (int t = 0, d = 0; t < size; t++) { cnt[t] = (int)(nout[t] * elsize); dsp[t] = d; d += cnt[t]; }
    unsigned char *all = rank == 0 ? malloc(ntot * elsize + 1) : NULL;
    MPI_Gatherv(out, (int)(nout[rank] * elsize), MPI_BYTE, all, cnt, dsp, MPI_BYTE, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        FILE *g = fopen(fout, "wb");
        if (!g) MPI_Abort(MPI_COMM_WORLD, 5);
        fwrite(all, elsize, ntot, g);
        fclose(g);
        qsort(times, repeat, sizeof(double), cmpd);
        printf("median_s %.9f\n", repeat > 0 ? times[repeat / 2] : 0.0);
    }
    (void)ooff;
    MPI_Finalize();
    return 0;
}
