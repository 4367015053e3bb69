/*
 * mpsort.h -- source-compatible replacement of the reference header
 * (mpsort.h:1-56 of MP-sort).  Callers written against the reference compile
 * unchanged and link against libmpsort_gpu.so (and, for the MPI-shaped
 * symbols, libmpsort_mpi_gpu.so).
 *
 *   radix_sort            mpsort.h:1-4      stable sort by radix; one virtual
 *                                           rank set (mpsort_gpu_set_vranks)
 *   mpsort_omp            mpsort.h:8-11     same result as radix_sort
 *   mpsort_mpi_*_options  mpsort.h:17-24
 *   mpsort_mpi_impl       mpsort.h:26-34    communicator of size 1; the P
 *   mpsort_mpi_newarray_impl mpsort.h:36-47 virtual ranks come from
 *                                           mpsort_gpu_set_vranks()
 *   mpsort_mpi_report_last_run mpsort.h:49
 */
#ifndef MPSORT_H_GPU_DROPIN
#define MPSORT_H_GPU_DROPIN
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

void radix_sort(void * base, size_t nmemb, size_t size,
        void (*radix)(const void * ptr, void * radix, void * arg),
        size_t rsize,
        void * arg);

#ifdef _OPENMP
void mpsort_omp(void * base, size_t nmemb, size_t size,
        void (*radix)(const void * ptr, void * radix, void * arg),
        size_t rsize,
        void * arg);
#endif

#define MPSORT_DISABLE_SPARSE_ALLTOALLV (1 << 1)
#define MPSORT_DISABLE_GATHER_SORT (1 << 3)
#define MPSORT_REQUIRE_GATHER_SORT (1 << 4)
#define MPSORT_REQUIRE_SPARSE_ALLTOALLV (1 << 6)

void mpsort_mpi_set_options(int options);
int mpsort_mpi_has_options(int options);
void mpsort_mpi_unset_options(int options);
void mpsort_mpi_report_last_run();

#ifdef MPI_VERSION
void mpsort_mpi_impl(void * base, size_t nmemb, size_t elsize,
        void (*radix)(const void * ptr, void * radix, void * arg),
        size_t rsize,
        void * arg, MPI_Comm comm,
        const int line, const char * file);

#define mpsort_mpi(base, nmemb, elsize, radix, rsize, arg, comm) \
    mpsort_mpi_impl(base, nmemb, elsize, radix, rsize, arg, comm, \
    __LINE__, __FILE__)

void mpsort_mpi_newarray_impl(void * base, size_t nmemb,
        void * out, size_t outnmemb,
        size_t elsize,
        void (*radix)(const void * ptr, void * radix, void * arg),
        size_t rsize,
        void * arg, MPI_Comm comm,
        const int line, const char * file);

#define mpsort_mpi_newarray(base, nmemb, out, outnmemb, elsize, \
    radix, rsize, arg, comm) \
    mpsort_mpi_newarray_impl(base, nmemb, out, outnmemb, elsize, \
    radix, rsize, arg, comm, __LINE__, __FILE__)
#endif

#ifdef __cplusplus
}
#endif
#endif
