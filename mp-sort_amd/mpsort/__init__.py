"""mpsort -- the MP-sort Python interface, sorting on one MI355X.

The public functions keep the names, arguments and results of MP-sort's
package (mpsort/__init__.py:23-204): ``sort``, ``permute``, ``take``,
``histogram``, ``globalrange`` and ``globalindices``.  ``comm`` is a
:class:`VirtualComm` (or None): one process holds the whole array and the
GPU sort splits it into virtual ranks, which stand for the reference's MPI
ranks.  The output never depends on that split, because the sort is globally
stable.

How the array forms work here.  A key given as an array (not a field name)
is sorted as ``{key, position}`` records on the GPU; the positions that come
back are the stable order, and the payload is gathered by them on the host.
``take`` and ``permute`` sort the requested indices once, read the source in
that (increasing) order and scatter the values back to the requesting slots.
"""
from .version import __version__
from .binding import sort as _sort_field, VirtualComm
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


def _elements(a):
    """(ndarray, item shape): an array as a sequence of items.

    A flat iterator stands for every scalar of its array, in C order; an
    ndarray's items are its rows (shape[1:] per item)."""
    if isinstance(a, numpy.flatiter):
        return numpy.asarray(a), ()
    a = numpy.asarray(a)
    return a, a.shape[1:]


def guess_dtype(array):
    """(dtype, item shape) of the items of an array or flat iterator."""
    arr, shape = _elements(array)
    return arr.dtype, shape


def _store(out, values):
    """Write values (one per item) into out, an ndarray or a flat iterator."""
    if isinstance(out, numpy.flatiter):
        out[:] = values.reshape(-1)
    else:
        out[...] = values


def _index_width(nkey_bytes, n):
    """Bytes of the position field and of padding after a key of nkey_bytes.

    The record {key, position, pad} is sized 8, 16 or 32 bytes when it can be,
    so the GPU sorts it without packing."""
    need_wide = n >= (1 << 32)
    for size in (8, 16, 32):
        room = size - nkey_bytes
        if room >= 8:
            return 8, room - 8
        if room >= 4 and not need_wide:
            return 4, room - 4
    return (8 if need_wide else 4), 0


def _stable_order(keys, item_shape, comm, tuning):
    """Positions 0..n-1 of `keys` in stable sorted order (GPU sort)."""
    n = len(keys)
    kdt = numpy.dtype((keys.dtype, item_shape)) if item_shape else keys.dtype
    width, pad = _index_width(kdt.itemsize, n)
    fields = [('key', kdt), ('pos', 'u%d' % width)]
    if pad:
        fields.append(('pad', 'V%d' % pad))
    recs = numpy.empty(n, dtype=fields)
    recs['key'] = keys
    recs['pos'] = numpy.arange(n, dtype=recs.dtype['pos'])
    _sort_field(recs, 'key', None, comm=comm, tuning=tuning)
    return recs['pos']


def sort(source, orderby=None, out=None, comm=None, tuning=[]):
    """Sort `source` by the key `orderby` into `out` (in place when out is None).

    orderby: a field name of source, an array of keys (one per item of
    source, on the same partition), or None to sort source by itself.  Keys
    are u8, i8, u4 or i4, scalars or 1-d vectors whose later elements are
    more significant.  out: same total length and item size as source.
    tuning: option names ('DISABLE_SPARSE_ALLTOALLV', 'DISABLE_GATHER_SORT',
    'REQUIRE_GATHER_SORT', 'REQUIRE_SPARSE_ALLTOALLV'); the SPARSE ones pick
    the exchange's run table (every rank pair, or the non-empty ones only),
    the GATHER ones are recorded; none changes the result.  source, orderby
    and out may be flat iterators.
    Returns out (None for a field-name key, as the reference).
    """
    if isinstance(orderby, (str, bytes)):
        return _sort_field(source, orderby, out, comm=comm, tuning=tuning)
    data, data_shape = _elements(source)
    if orderby is None:
        keys, key_shape = data, data_shape
    else:
        keys, key_shape = _elements(orderby)
    if len(keys) != len(data):
        raise ValueError("orderby has %d items, source has %d" % (len(keys), len(data)))
    target = source if out is None else out
    tdata, _ = _elements(target)
    if len(tdata) != len(data):
        raise ValueError("total size of array changed %d != %d" % (len(data), len(tdata)))
    if len(data) == 0:
        return target
    order = _stable_order(keys, key_shape, comm, tuning)
    # fancy indexing copies, so an in-place sort reads the old items
    _store(target, data[order])
    return target


def globalrange(array, comm):
    """[start, end) of this process's chunk in the global array."""
    sizes = comm.allgather(len(array))
    start = int(sum(sizes[:comm.rank]))
    return start, start + len(array)


def globalindices(array, comm):
    """Global indices of this process's items (int64 above 2^30 items in
    total, int32 otherwise)."""
    start, end = globalrange(array, comm)
    total = comm.allreduce(len(array))
    return numpy.arange(start, end, dtype=numpy.int64 if total > (1 << 30) else numpy.int32)


def histogram(array, bins, comm, right=False):
    """Counts of `array` in the len(bins)+1 intervals cut by the increasing
    edges `bins` (numpy.digitize convention), summed over the communicator."""
    nslot = len(bins) + 1
    values = numpy.asarray(array)
    if values.size == 0:
        local = numpy.zeros(nslot, dtype=numpy.intp)
    else:
        local = numpy.bincount(numpy.digitize(values, bins, right), minlength=nslot)
    return comm.allreduce(local)


def take(source, argindex, comm, out=None):
    """source[argindex], distributed as argindex.  argindex need be neither
    a permutation nor unique; its values are global indices into source."""
    data, shape = _elements(source)
    want, want_shape = _elements(argindex)
    if want_shape:
        raise ValueError("argindex must be one index per item")
    if out is None:
        out = numpy.empty(len(want), dtype=(data.dtype, shape) if shape else data.dtype)
    if len(want) == 0:
        return out
    start, end = globalrange(data, comm)
    # requests in increasing index order: the source is read front to back
    # (argindex has its own length, so only the comm's rank count carries over)
    sorter = VirtualComm(vranks=comm.vranks) if isinstance(comm, VirtualComm) else comm
    order = _stable_order(want, (), sorter, [])
    wanted = want[order].astype(numpy.int64) - start
    if wanted.size and (wanted.min() < 0 or wanted.max() >= len(data)):
        raise IndexError("argindex holds indices outside [0, %d)" % (end - start))
    values = data[wanted]
    result = numpy.empty((len(want),) + shape, dtype=data.dtype)
    result[order] = values
    _store(out, result)
    return out


def permute(source, argindex, comm, out=None):
    """source[argindex] for a permutation argindex of the global indices,
    distributed as argindex."""
    if comm.allreduce(len(source)) != comm.allreduce(len(argindex)):
        raise ValueError("Global size of source and argindex is different")
    return take(source, argindex, comm, out=out)
