/*
 * A C caller written against the reference API (mpsort.h with MPI), linked
 * unchanged against libmpsort_mpi_gpu.so + libmpsort_gpu.so.
 *
 *   mpi_sort_example NBASE SEED SHIFT INPLACE OUTFILE
 *
 * Rank r holds NBASE + 37 r records {u64 key, u64 index}; record g (global
 * index, ranks in order) has key = splitmix64(SEED + g) >> SHIFT (a large
 * SHIFT gives many equal keys) and index g.  The records are sorted by key
 * with mpsort_mpi_newarray into an even split of the total (or in place with
 * mpsort_mpi when INPLACE is 1), and rank 0 writes every rank's output, in
 * rank order, to OUTFILE.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mpsort.h"

struct record {
    uint64_t key;
    uint64_t index;
};

static uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void radix(const void *ptr, void *out, void *arg)
{
    (void)arg;
    memcpy(out, ptr, 8); /* the u64 key is its own radix */
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc != 6) {
        if (rank == 0) fprintf(stderr, "usage: %s NBASE SEED SHIFT INPLACE OUTFILE\n", argv[0]);
        MPI_Finalize();
        return 2;
    }
    const long nbase = atol(argv[1]);
    const uint64_t seed = strtoull(argv[2], NULL, 0);
    const int shift = atoi(argv[3]);
    const int inplace = atoi(argv[4]);

    long first = 0, total = 0;
    for (int r = 0; r < size; r++) {
        if (r < rank) first += nbase + 37L * r;
        total += nbase + 37L * r;
    }
    const size_t nin = (size_t)(nbase + 37L * rank);
    const size_t nout = inplace ? nin : (size_t)(total * (rank + 1) / size - total * rank / size);
    struct record *in = malloc(sizeof(struct record) * (nin + 1));
    struct record *out = inplace ? in : malloc(sizeof(struct record) * (nout + 1));
    for (size_t i = 0; i < nin; i++) {
        const uint64_t g = (uint64_t)first + i;
        in[i].key = splitmix64(seed + g) >> shift;
        in[i].index = g;
    }
    if (inplace)
        mpsort_mpi(in, nin, sizeof(struct record), radix, 8, NULL, MPI_COMM_WORLD);
    else
        mpsort_mpi_newarray(in, nin, out, nout, sizeof(struct record), radix, 8, NULL, MPI_COMM_WORLD);

    int *cnt = NULL, *displ = NULL;
    struct record *all = NULL;
    if (rank == 0) {
        cnt = malloc(sizeof(int) * 2 * size);
        displ = cnt + size;
        all = malloc(sizeof(struct record) * (size_t)(total + 1));
    }
    const int mine = (int)(nout * sizeof(struct record));
    MPI_Gather(&mine, 1, MPI_INT, cnt, 1, MPI_INT, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        int d = 0;
        for (int r = 0; r < size; r++) {
            displ[r] = d;
            d += cnt[r];
        }
    }
    MPI_Gatherv(out, mine, MPI_BYTE, all, cnt, displ, MPI_BYTE, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        FILE *f = fopen(argv[5], "wb");
        if (!f || fwrite(all, sizeof(struct record), (size_t)total, f) != (size_t)total) {
            fprintf(stderr, "cannot write %s\n", argv[5]);
            MPI_Abort(MPI_COMM_WORLD, 1);
        }
        fclose(f);
        printf("ranks %d records %ld\n", size, total);
    }
    MPI_Finalize();
    return 0;
}
