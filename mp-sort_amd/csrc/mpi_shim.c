/*
 * mpi_shim.c -- the MPI-shaped entry points of mpsort.h (reference
 * mpsort.h:26-47, mpsort-mpi.c:121-331) on top of the one-GPU histogram sort
 * of libmpsort_gpu.so.  Built with the MPI compiler wrapper into
 * libmpsort_mpi_gpu.so.
 *
 *   communicator of size 1: the records are split evenly over the virtual
 *     ranks set with mpsort_gpu_set_vranks() (or MPSORT_GPU_VRANKS), as
 *     mpsort_omp splits over its threads (mpsort-omp.c:242-243);
 *   communicator of size N > 1: rank 0 gathers every rank's records, sorts
 *     them on its GPU with one virtual rank per MPI rank (the ranks' own
 *     input and output counts) and scatters the result, so each rank ends
 *     with exactly the records the reference would leave it.
 *
 * Fatal conditions print the caller's file:line and abort the communicator,
 * as the reference does (mpsort-mpi.c:74,185,231,329).
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/mpsort.h"
#include "../../include/mpsort_gpu.h"
#include "../../include/mp-mpiu.h"

/* MPIU_Set_verbose_malloc (mp-mpiu.h:11): the host allocation hook prints
 * every allocation and free with the caller's rank */
static void *verbose_malloc(const char *name, size_t size, const char *file, const int line, void *userdata)
{
    int rank = 0;
    MPI_Comm_rank((MPI_Comm)(intptr_t)userdata, &rank);
    void *p = malloc(size);
    fprintf(stderr, "MPIU_Malloc: T%04d %16p : %s size = %zu, %s:%d\n", rank, p, name, size, file, line);
    return p;
}

static void verbose_free(void *ptr, const char *file, const int line, void *userdata)
{
    int rank = 0;
    MPI_Comm_rank((MPI_Comm)(intptr_t)userdata, &rank);
    fprintf(stderr, "MPIU_Free: T%04d %16p : %s:%d\n", rank, ptr, file, line);
    free(ptr);
}

void MPIU_Set_verbose_malloc(MPI_Comm comm)
{
    mpiu_set_malloc(verbose_malloc, verbose_free, (void *)(intptr_t)comm);
}

static void fatal(MPI_Comm comm, int line, const char *file, const char *what, const char *detail)
{
    fprintf(stderr, "%s:%d: %s%s%s\n", file, line, what, detail ? ": " : "", detail ? detail : "");
    fflush(stderr);
    MPI_Abort(comm, -1);
}

void mpsort_mpi_newarray_impl(void *base, size_t nmemb, void *out, size_t outnmemb, size_t elsize,
                              void (*radix)(const void *ptr, void *radix, void *arg), size_t rsize,
                              void *arg, MPI_Comm comm, const int line, const char *file)
{
    int size = 1, rank = 0;
    MPI_Comm_size(comm, &size);
    MPI_Comm_rank(comm, &rank);
    if (size == 1) {
        const int P = mpsort_gpu_get_vranks();
        int64_t *nin = malloc(sizeof(int64_t) * 2 * (size_t)P);
        int64_t *nout = nin + P;
        for (int t = 0; t < P; t++) {
            nin[t] = (int64_t)(nmemb * (size_t)(t + 1) / (size_t)P - nmemb * (size_t)t / (size_t)P);
            nout[t] = (int64_t)(outnmemb * (size_t)(t + 1) / (size_t)P - outnmemb * (size_t)t / (size_t)P);
        }
        const int rc = mpsort_gpu_newarray_radix(base, nin, P, out, nout, elsize, radix, rsize, arg, 0);
        free(nin);
        if (rc) fatal(comm, line, file, "mpsort_mpi_newarray", mpsort_gpu_last_error());
        return;
    }

    /* every rank learns every rank's input and output count */
    int64_t mine[2] = {(int64_t)nmemb, (int64_t)outnmemb};
    int64_t *all = malloc(sizeof(int64_t) * 2 * (size_t)size);
    MPI_Allgather(mine, 2, MPI_INT64_T, all, 2, MPI_INT64_T, comm);
    int64_t tin = 0, tout = 0;
    for (int r = 0; r < size; r++) {
        tin += all[2 * r];
        tout += all[2 * r + 1];
        if (all[2 * r] > INT32_MAX || all[2 * r + 1] > INT32_MAX)
            fatal(comm, line, file, "mpsort_mpi_newarray", "more than 2^31-1 records on one rank");
    }
    if (tin != tout) fatal(comm, line, file, "total number of items in the input and output differ", NULL);

    MPI_Datatype rec;
    MPI_Type_contiguous((int)elsize, MPI_BYTE, &rec);
    MPI_Type_commit(&rec);
    int *cin = NULL, *din = NULL, *cout = NULL, *dout = NULL;
    char *gin = NULL, *gout = NULL;
    int64_t *nin = NULL, *nout = NULL;
    if (rank == 0) {
        cin = malloc(sizeof(int) * 4 * (size_t)size);
        din = cin + size;
        cout = din + size;
        dout = cout + size;
        nin = malloc(sizeof(int64_t) * 2 * (size_t)size);
        nout = nin + size;
        int64_t a = 0, b = 0;
        for (int r = 0; r < size; r++) {
            nin[r] = all[2 * r];
            nout[r] = all[2 * r + 1];
            cin[r] = (int)nin[r];
            cout[r] = (int)nout[r];
            if (a > INT32_MAX || b > INT32_MAX)
                fatal(comm, line, file, "mpsort_mpi_newarray", "more than 2^31-1 records in total");
            din[r] = (int)a;
            dout[r] = (int)b;
            a += nin[r];
            b += nout[r];
        }
        gin = mpiu_malloc("mpsort_mpi:gather", elsize * (size_t)(tin > 0 ? tin : 1), __FILE__, __LINE__);
        gout = mpiu_malloc("mpsort_mpi:scatter", elsize * (size_t)(tin > 0 ? tin : 1), __FILE__, __LINE__);
        if (!gin || !gout) fatal(comm, line, file, "mpsort_mpi_newarray", "host allocation failed");
    }
    MPI_Gatherv(base, (int)nmemb, rec, gin, cin, din, rec, 0, comm);
    int rc = 0;
    if (rank == 0) rc = mpsort_gpu_newarray_radix(gin, nin, size, gout, nout, elsize, radix, rsize, arg, 0);
    MPI_Bcast(&rc, 1, MPI_INT, 0, comm);
    if (rc) {
        if (rank == 0) fatal(comm, line, file, "mpsort_mpi_newarray", mpsort_gpu_last_error());
        MPI_Barrier(comm);
    }
    MPI_Scatterv(gout, cout, dout, rec, out, (int)outnmemb, rec, 0, comm);
    MPI_Type_free(&rec);
    free(all);
    free(cin);
    free(nin);
    mpiu_free(gin, __FILE__, __LINE__);
    mpiu_free(gout, __FILE__, __LINE__);
}

void mpsort_mpi_impl(void *base, size_t nmemb, size_t elsize,
                     void (*radix)(const void *ptr, void *radix, void *arg), size_t rsize, void *arg,
                     MPI_Comm comm, const int line, const char *file)
{
    mpsort_mpi_newarray_impl(base, nmemb, base, nmemb, elsize, radix, rsize, arg, comm, line, file);
}
