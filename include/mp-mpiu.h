/*
 * mp-mpiu.h -- the allocator hook of the reference's MPI utility header
 * (mp-mpiu.h:1-21 of MP-sort), so that callers which install their own
 * allocator -- the reference's Cython binding does, at import
 * (mpsort/binding.pyx:30-41) -- compile and link unchanged.
 *
 *   mpiu_set_malloc / MPIU_SetMalloc   mp-mpiu.h:10,16   libmpsort_gpu.so
 *   mpiu_malloc / mpiu_free            mp-mpiu.h:7-8     libmpsort_gpu.so
 *   MPIU_Set_verbose_malloc            mp-mpiu.h:11      libmpsort_mpi_gpu.so
 *                                                        (needs MPI_Comm)
 *
 * The hook serves the library's large HOST temporaries, as the reference's
 * serves its exchange and gather buffers (mpsort-mpi.c:278-279,572): the
 * materialised radix of the callback form and the gather/scatter buffers of
 * the MPI entry points.  Device memory has its own hook,
 * mpsort_gpu_set_allocator (mpsort_gpu.h).  The default is malloc/free.
 *
 * The reference's transport helpers declared beside the hook (MPIU_Alltoallv,
 * MPIU_GetLoc, MPIU_Gather, MPIU_Scatter, MPIU_Segmenter) are not part of
 * the sort's interface; on one GPU the exchange is a device pass, and they
 * are not provided.
 */
#ifndef MP_MPIU_H_GPU_DROPIN
#define MP_MPIU_H_GPU_DROPIN
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void * (*mpiu_malloc_func)(const char * name, size_t size, const char * file, const int line, void * userdata);
typedef void (*mpiu_free_func)(void * ptr, const char * file, const int line, void * userdata);

void * mpiu_malloc(const char * name, size_t size, const char * file, const int line);
void mpiu_free(void * ptr, const char * file, const int line);

/* NULL for either function restores malloc/free. */
void mpiu_set_malloc(mpiu_malloc_func malloc, mpiu_free_func free, void * userdata);

#ifdef MPI_VERSION
/* every allocation and free is printed to stderr with the caller's rank */
void MPIU_Set_verbose_malloc(MPI_Comm comm);
#endif

#define MPIU_SetMalloc mpiu_set_malloc
#define MPIU_MallocT(name, type, nmemb) MPIU_Malloc(name, sizeof(type), nmemb)
#define MPIU_Malloc(name, elsize, nmemb) mpiu_malloc(name, ((size_t)(elsize)) * nmemb, __FILE__, __LINE__)
#define MPIU_Free(ptr) mpiu_free(ptr, __FILE__, __LINE__)

#ifdef __cplusplus
}
#endif
#endif
