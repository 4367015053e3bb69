"""Batched segmented sort: many independent stable sorts in one plan (the
reference's counterpart is a loop of radix_sort, radixsort.c:35-44, one call
per segment)."""
import ctypes

import numpy

from . import _lib
from ._lib import Stats, check, lib
from .binding import _key_desc


class SegmentPlan(object):
    """Groups of segments on the GPU, sorted by one call of :meth:`run`.

    Each group is a device buffer of records of one dtype holding its
    segments back to back.  ``graph=True`` captures the whole plan in a
    hipGraph on the first run and replays it afterwards; ``tune=True`` picks
    the tile shape, the in-LDS threshold and the workgroup shape of the
    in-LDS segments by timing them on the first run.  ``small_form`` sets that
    shape without tuning: 0 (1024 threads x 9 records, one workgroup per CU),
    1 (512 x 9, two per CU) or 2 (512 x 6, three per CU)."""

    def __init__(self, graph=False, tune=False, small_form=0, each=False):
        flags = (_lib.SEGPLAN_GRAPH if graph else 0) | (_lib.SEGPLAN_TUNE if tune else 0)
        flags |= {0: 0, 1: _lib.SEGPLAN_SMALL_512x9, 2: _lib.SEGPLAN_SMALL_512x6}[small_form]
        # each: every segment by launches of its own, one after the other (the
        # per-segment loop the batched form replaces)
        flags |= _lib.SEGPLAN_EACH if each else 0
        h = ctypes.c_void_p()
        check(lib.mpsort_gpu_segplan_create(ctypes.byref(h), flags))
        self._h = h
        self._keep = []

    def add(self, d_in, d_out, seg_sizes, dtype, orderby=None):
        """Add a group: device pointers d_in / d_out (may be equal), sizes of
        the segments in order, the record dtype and the key field."""
        dtype = numpy.dtype(dtype)
        sizes = numpy.ascontiguousarray(seg_sizes, dtype=numpy.int64)
        key = _key_desc(dtype, orderby)
        self._keep.append((sizes, key))
        check(lib.mpsort_gpu_segplan_add(self._h, d_in, d_out,
                                         sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(sizes),
                                         dtype.itemsize, ctypes.byref(key)))

    def run(self):
        """Sort every group; returns the device time in ms."""
        st = Stats()
        check(lib.mpsort_gpu_segplan_run(self._h, ctypes.byref(st)))
        return st.t_total_ms

    def info(self, group=0):
        v = [ctypes.c_int() for _ in range(5)]
        check(lib.mpsort_gpu_segplan_info(self._h, group, *[ctypes.byref(x) for x in v]))
        d = dict(zip(("small_segments", "large_segments", "block", "items", "threshold"), [x.value for x in v]))
        f, w = ctypes.c_int(), ctypes.c_int()
        check(lib.mpsort_gpu_segplan_shape(self._h, group, ctypes.byref(f), ctypes.byref(w)))
        d["small_form"], d["small_waves_per_cu"] = f.value, w.value
        return d

    def close(self):
        if self._h:
            lib.mpsort_gpu_segplan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_segments(data, seg_sizes, orderby=None, graph=False, tune=False, runs=1):
    """Sort the segments of a host array independently (stable, by the key
    field); returns the sorted copy."""
    data = numpy.ascontiguousarray(data)
    nb = data.nbytes
    d = lib.mpsort_gpu_dmalloc(max(nb, 1))
    if not d:
        raise MemoryError("device allocation failed")
    try:
        check(lib.mpsort_gpu_memcpy_h2d(d, data.ctypes.data, nb))
        plan = SegmentPlan(graph=graph, tune=tune)
        try:
            plan.add(d, d, seg_sizes, data.dtype, orderby)
            for _ in range(runs):
                if runs > 1:
                    check(lib.mpsort_gpu_memcpy_h2d(d, data.ctypes.data, nb))
                plan.run()
        finally:
            plan.close()
        out = numpy.empty_like(data)
        check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, d, nb))
    finally:
        lib.mpsort_gpu_dfree(d)
    return out
