"""Binding of ``mpsort.sort`` to the GPU library: the counterpart of
binding.pyx (MP-sort) with the per-record radix callback replaced by the key
descriptor that radix_data_init (binding.pyx:50-79) derives from the dtype.
"""
import ctypes

import numpy

from . import _lib
from ._lib import KeyDesc, check, lib


class VirtualComm(object):
    """One-process communicator for the GPU sort.

    Collectives behave as on a communicator of size 1 (``allreduce(x) == x``,
    ``allgather(x) == [x]``): one process holds the whole array.  The sort
    itself splits that array into ``vranks`` virtual ranks, the reference's
    MPI ranks, with ``in_sizes`` / ``out_sizes`` records each (even split, as
    mpsort-omp.c:242-243, when not given).  The output does not depend on the
    partition: the reference's sort is globally stable.
    """

    rank = 0
    size = 1

    def __init__(self, vranks=None, in_sizes=None, out_sizes=None):
        if in_sizes is not None:
            vranks = len(in_sizes)
        if vranks is None:
            vranks = int(lib.mpsort_gpu_get_vranks())
        self.vranks = int(vranks)
        if self.vranks < 1:
            raise ValueError("vranks must be >= 1")
        self.in_sizes = None if in_sizes is None else [int(x) for x in in_sizes]
        self.out_sizes = None if out_sizes is None else [int(x) for x in out_sizes]
        if self.out_sizes is not None and len(self.out_sizes) != self.vranks:
            raise ValueError("out_sizes must have one entry per virtual rank")

    # mpi4py-shaped collectives of a size-1 communicator
    def allreduce(self, x, op=None):
        return x

    def allgather(self, x):
        return [x]

    def bcast(self, x, root=0):
        return x

    def gather(self, x, root=0):
        return [x]

    def scatter(self, x, root=0):
        return x[0]

    def barrier(self):
        pass

    Barrier = barrier

    def partition(self, n, sizes):
        if sizes is None:
            P = self.vranks
            return [n * (t + 1) // P - n * t // P for t in range(P)]
        if sum(sizes) != n:
            raise ValueError("partition sizes %s do not add up to %d" % (sizes, n))
        return list(sizes)


COMM_WORLD = None


def _key_desc(dtype, radixkey):
    """radix_data_init (binding.pyx:50-79)."""
    if radixkey is not None:
        radixdtype, offset = dtype.fields[radixkey][:2]
    else:
        radixdtype, offset = dtype, 0
    if len(radixdtype.shape) == 0:
        nmemb = 1
    elif len(radixdtype.shape) == 1:
        nmemb = radixdtype.shape[0]
    else:
        raise ValueError("data[%s] is not 1d nor 2d" % (radixkey,))
    base = radixdtype.base
    if base == numpy.dtype('u8'):
        width, signed = 8, 0
    elif base == numpy.dtype('i8'):
        width, signed = 8, 1
    elif base == numpy.dtype('u4'):
        width, signed = 4, 0
    elif base == numpy.dtype('i4'):
        width, signed = 4, 1
    else:
        raise TypeError("data[%s] is not u8 or i8" % (radixkey,))
    return KeyDesc(offset=int(offset), width=width, is_signed=signed, nwords=int(nmemb), reserved=0)


_TUNING = {
    'DISABLE_SPARSE_ALLTOALLV': _lib.MPSORT_DISABLE_SPARSE_ALLTOALLV,
    'DISABLE_GATHER_SORT': _lib.MPSORT_DISABLE_GATHER_SORT,
    'REQUIRE_GATHER_SORT': _lib.MPSORT_REQUIRE_GATHER_SORT,
    'REQUIRE_SPARSE_ALLTOALLV': _lib.MPSORT_REQUIRE_SPARSE_ALLTOALLV,
}


def _as_comm(comm):
    if comm is None:
        return VirtualComm()
    if isinstance(comm, VirtualComm):
        return comm
    raise ValueError("only mpsort.VirtualComm objects (or None) are supported")


def sort(data, orderby=None, out=None, comm=None, tuning=[], flags=0):
    """Sort ``data`` by ``data[orderby]`` into ``out`` (binding.pyx:123-209).

    data[orderby] must be of u8/i8/u4/i4 type, scalar or a 1-d subarray in
    which the later elements are more significant.  out=None sorts in place.
    """
    if not isinstance(data, numpy.ndarray):
        raise TypeError("data must be a numpy.ndarray")
    # assert you can access the orderby columns
    if orderby is not None:
        data[orderby]
    if not data.flags['C_CONTIGUOUS']:
        raise ValueError("data must be C_CONTIGUOUS")
    if out is None:
        out = data
    if not isinstance(out, numpy.ndarray):
        raise TypeError("out must be a numpy.ndarray")
    if not out.flags['C_CONTIGUOUS']:
        raise ValueError("out must be C_CONTIGUOUS")
    comm = _as_comm(comm)
    Ntot = comm.allreduce(len(data))
    Ntotout = comm.allreduce(len(out))
    if Ntot != Ntotout:
        raise ValueError("total size of array changed %d != %d" % (Ntot, Ntotout))
    if data.dtype.itemsize != out.dtype.itemsize:
        raise ValueError("item size mismatch")
    key = _key_desc(data.dtype, orderby)

    # binding.pyx:195-204: every call starts from no options
    lib.mpsort_mpi_unset_options(-1)
    for name, bit in _TUNING.items():
        if name in tuning:
            lib.mpsort_mpi_set_options(bit)

    nin = comm.partition(len(data), comm.in_sizes)
    nout = comm.partition(len(out), comm.out_sizes)
    P = len(nin)
    if len(nout) != P:
        raise ValueError("input and output partitions have different rank counts")
    nin_a = (ctypes.c_int64 * P)(*nin)
    nout_a = (ctypes.c_int64 * P)(*nout)
    if Ntot == 0:
        return
    src = data
    if out is not data and numpy.shares_memory(data, out):
        src = data.copy()
    rc = lib.mpsort_gpu_newarray(src.ctypes.data, nin_a, P, out.ctypes.data, nout_a,
                                 data.dtype.itemsize, ctypes.byref(key), int(flags))
    check(rc)
