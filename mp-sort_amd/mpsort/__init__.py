"""mpsort -- MI355X histogram sort behind the MP-sort Python API.

``sort``, ``permute``, ``take``, ``histogram``, ``globalrange`` and
``globalindices`` keep the signatures and semantics of the reference package
(mpsort/__init__.py:23-204 of MP-sort).  The sort runs on the GPU through
libmpsort_gpu.so; ``comm`` is a :class:`VirtualComm` (or None) whose virtual
ranks stand for the reference's MPI ranks.
"""
from .version import __version__
from .binding import sort as _sort, VirtualComm
from . import _lib

import numpy

__all__ = ["sort", "permute", "take", "histogram", "globalrange", "globalindices",
           "VirtualComm", "set_virtual_ranks", "last_stats", "__version__"]


def set_virtual_ranks(P):
    """Number of virtual ranks used when comm is None (default 1)."""
    _lib.lib.mpsort_gpu_set_vranks(int(P))


def last_stats():
    """Phase timers and pass counts of the last sort (mpsort_gpu_stats)."""
    return _lib.last_stats().as_dict()


def sort(source, orderby=None, out=None, comm=None, tuning=[]):
    """
        Sort source array with orderby as the key.
        Store result to out.

        Parameters
        ----------
        source : array, 1d, distributed

        orderby : array, 1d, distributed or string.
            Only integer types are supported.
            must be on the same partition as that of source.
            If orderby is string, it refers to the field in source.

        out : array, 1d distributed
            the total length must be the same as source.
            the itemsize must be the same as source
            if None, the sort is in-place.

        tuning: list of strings
            'DISABLE_SPARSE_ALLTOALLV'
            'DISABLE_GATHER_SORT'
            'REQUIRE_GATHER_SORT'
            'REQUIRE_SPARSE_ALLTOALLV'

        Returns
        -------
            out

        Remarks
        -------
            source, orderby, out can be flatiter at the same time.
    """
    key = orderby
    if isinstance(key, (str, bytes)):
        return _sort(source, key, out, comm=comm, tuning=tuning)

    # mpsort/__init__.py:62-83: pack data (and key) into one struct array
    if key is None:
        D, I = 'DD'
        data1 = numpy.empty(len(source), dtype=[('D', guess_dtype(source))])
        data1['D'][...] = source
    else:
        D, I = 'DI'
        data1 = numpy.empty(len(source), dtype=[('D', guess_dtype(source)),
                                                ('I', guess_dtype(key))])
        data1['D'][...] = source
        data1['I'][...] = key

    if out is None:
        out = source
        _sort(data1, orderby=I, comm=comm, tuning=tuning)
        out[...] = data1[D][...]
    else:
        data2 = numpy.empty(len(out), dtype=data1.dtype)
        _sort(data1, orderby=I, out=data2, comm=comm, tuning=tuning)
        out[...] = data2[D][...]
    return out


def globalrange(array, comm):
    """The start and end of local chunk in the global array."""
    s = comm.allgather(len(array))
    start = sum(s[:comm.rank])
    end = start + s[comm.rank]
    return (start, end)


def globalindices(array, comm):
    """Construct a list of indices in the global array."""
    start, end = globalrange(array, comm)
    globalsize = comm.bcast(end, root=comm.size - 1)
    if globalsize > 1024 * 1024 * 1024:
        dtype = 'i8'
    else:
        dtype = 'i4'
    return numpy.arange(start, end, dtype=dtype)


def guess_dtype(array):
    if isinstance(array, numpy.flatiter):
        array = array.base
        return array.dtype, ()
    return array.dtype, array.shape[1:]


def permute(source, argindex, comm, out=None):
    """Permute a distributed array: source[argindex], on argindex's partition."""
    source_size = comm.allreduce(len(source))
    argindex_size = comm.allreduce(len(argindex))
    if source_size != argindex_size:
        raise ValueError("Global size of source and argindex is different")
    if out is None:
        out = numpy.empty(len(argindex), guess_dtype(source))
    originind = globalindices(argindex, comm)
    originind2 = numpy.empty(len(source), dtype=originind.dtype)
    sort(originind, orderby=argindex, out=originind2, comm=comm)
    sort(source, orderby=originind2, out=out, comm=comm)
    return out


def histogram(array, bins, comm, right=False):
    """Histogram of a distributed array over collective bin edges."""
    if len(array) == 0:
        originrank = []
    else:
        originrank = numpy.digitize(array, bins, right)
    recv = numpy.bincount(originrank, minlength=len(bins) + 1)
    return comm.allreduce(recv)


def take(source, argindex, comm, out=None):
    """source[argindex] distributed as argindex (argindex need not be a
    permutation)."""
    start, end = globalrange(source, comm)
    bins = comm.allgather(end)
    h = histogram(argindex, bins, comm)
    nactive = h[comm.rank]
    if out is None:
        out = numpy.empty(len(argindex), guess_dtype(source))
    originind = globalindices(argindex, comm)
    myargindex = numpy.empty(nactive, dtype=guess_dtype(argindex))
    myoriginind = numpy.empty(nactive, dtype=originind.dtype)
    sort(originind, orderby=argindex, out=myoriginind, comm=comm)
    sort(argindex, orderby=argindex, out=myargindex, comm=comm)
    myresult = source[myargindex - start]
    sort(myresult, orderby=myoriginind, out=out, comm=comm)
    return out
