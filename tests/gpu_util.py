"""Helpers of the GPU parity tests: call the C ABI directly (ctypes)."""
import ctypes

import numpy

from mpsort import _lib
from mpsort._lib import KeyDesc, lib


def key_desc(dtype, field=None):
    dtype = numpy.dtype(dtype)
    if field is not None:
        kd, off = dtype.fields[field][:2]
    else:
        kd, off = dtype, 0
    nw = 1 if len(kd.shape) == 0 else kd.shape[0]
    w, s = {numpy.dtype('u8'): (8, 0), numpy.dtype('i8'): (8, 1),
            numpy.dtype('u4'): (4, 0), numpy.dtype('i4'): (4, 1)}[kd.base]
    return KeyDesc(offset=int(off), width=w, is_signed=s, nwords=int(nw), reserved=0)


def gpu_sort(a, nin, nout, field=None, flags=0, inplace=False):
    """mpsort_gpu_newarray on host memory; returns (out, stats)."""
    a = numpy.ascontiguousarray(a)
    P = len(nin)
    nin_a = (ctypes.c_int64 * P)(*[int(x) for x in nin])
    nout_a = (ctypes.c_int64 * P)(*[int(x) for x in nout])
    key = key_desc(a.dtype, field)
    if inplace:
        out = a
    else:
        out = numpy.zeros_like(a)
    rc = lib.mpsort_gpu_newarray(a.ctypes.data, nin_a, P, out.ctypes.data, nout_a, a.dtype.itemsize,
                                 ctypes.byref(key), flags)
    _lib.check(rc)
    return out, _lib.last_stats()


def gpu_sort_device(a, nin, nout, field=None, flags=0):
    """The same with device pointers (inputs resident in HBM)."""
    a = numpy.ascontiguousarray(a)
    nb = a.nbytes
    din = lib.mpsort_gpu_dmalloc(max(nb, 1))
    dout = lib.mpsort_gpu_dmalloc(max(nb, 1))
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(din, a.ctypes.data, nb))
        P = len(nin)
        nin_a = (ctypes.c_int64 * P)(*[int(x) for x in nin])
        nout_a = (ctypes.c_int64 * P)(*[int(x) for x in nout])
        key = key_desc(a.dtype, field)
        rc = lib.mpsort_gpu_newarray(din, nin_a, P, dout, nout_a, a.dtype.itemsize, ctypes.byref(key),
                                     flags | _lib.DEVICE_PTRS)
        _lib.check(rc)
        out = numpy.empty_like(a)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, dout, nb))
    finally:
        lib.mpsort_gpu_dfree(din)
        lib.mpsort_gpu_dfree(dout)
    return out, _lib.last_stats()


def last_counts(P, rsize):
    sp = numpy.zeros((max(P - 1, 1), rsize), numpy.uint8)
    pmin = numpy.zeros(rsize, numpy.uint8)
    pmax = numpy.zeros(rsize, numpy.uint8)
    C = numpy.zeros(P + 1, numpy.int64)
    clt = numpy.zeros((P, P + 1), numpy.int64)
    cle = numpy.zeros((P, P + 1), numpy.int64)
    myc = numpy.zeros((P, P + 1), numpy.int64)
    send = numpy.zeros((P, P), numpy.int64)
    rs = ctypes.c_int64(0)
    i64 = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    _lib.check(lib.mpsort_gpu_last_counts(sp.ctypes.data, pmin.ctypes.data, pmax.ctypes.data, i64(C),
                                          i64(clt), i64(cle), i64(myc), i64(send), ctypes.byref(rs)))
    return dict(P=sp[:P - 1], Pmin=pmin, Pmax=pmax, C=C, myCLT=clt, myCLE=cle, myC=myc, send=send,
                rsize=rs.value)


def even(n, P):
    return [n * (t + 1) // P - n * t // P for t in range(P)]
