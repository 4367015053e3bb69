"""The allocation-free entry point (mpsort_gpu_newarray_noalloc) against
mpsort_gpu_newarray on the same device inputs: output bytes, per-rank
counts and checksums, on two calls in a row (the second smaller), for the
direct and the packed record paths, and its size limit."""
import ctypes
import sys

import numpy
import pytest

from gpu_util import even, key_desc, last_counts
from mpsort import _lib
from mpsort._lib import lib

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _counts(v):
    return (ctypes.c_int64 * len(v))(*[int(x) for x in v])


def _run(fn, a, nin, nout, field, inplace=False, flags=0):
    """fn on device copies of a; returns (host output, stats, counts)"""
    nb = a.nbytes
    din = lib.mpsort_gpu_dmalloc(max(nb, 1))
    dout = din if inplace else lib.mpsort_gpu_dmalloc(max(nb, 1))
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(din, a.ctypes.data, nb))
        k = key_desc(a.dtype, field)
        rc = fn(din, _counts(nin), len(nin), dout, _counts(nout), a.dtype.itemsize, ctypes.byref(k),
                flags | _lib.DEVICE_PTRS)
        _lib.check(rc)
        out = numpy.empty_like(a)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, dout, nb))
        st = _lib.last_stats()
        P = len(nin)
        cnt = last_counts(P, 8 * k.nwords) if P > 1 and len(a) else None
    finally:
        lib.mpsort_gpu_dfree(din)
        if not inplace:
            lib.mpsort_gpu_dfree(dout)
    return out, st, cnt


def _same(a, nin, nout, field, inplace=False, flags=0):
    ref, st0, c0 = _run(lib.mpsort_gpu_newarray, a, nin, nout, field, inplace, flags)
    got, st1, c1 = _run(lib.mpsort_gpu_newarray_noalloc, a, nin, nout, field, inplace, flags)
    assert numpy.array_equal(got.view(numpy.uint8), ref.view(numpy.uint8)), "output differs"
    assert st1.checksum_in == st0.checksum_in and st1.checksum_out == st0.checksum_out
    assert st1.general_path == st0.general_path
    if c0 is not None:
        for name in ("P", "C", "myCLT", "myCLE", "myC", "send"):
            assert numpy.array_equal(c0[name], c1[name]), name
    return got


def rec16(n, seed, nkeys):
    g = numpy.random.default_rng(seed)
    a = numpy.zeros(n, dtype=[('key', 'u8'), ('idx', 'u8')])
    a['key'] = g.integers(0, nkeys, n)
    a['idx'] = numpy.arange(n)
    return a


def test_noalloc_equals_newarray_two_calls_second_smaller():
    """direct path, 16-byte records: 3,000,000 records on 64 ranks, then
    200,000 on 16 (the scratch keeps the first call's contents)"""
    a = rec16(3000000, 81, 1 << 40)
    note("noalloc direct n=%d P=64" % len(a))
    out = _same(a, even(len(a), 64), even(len(a), 64), 'key')
    assert numpy.array_equal(out, a[numpy.argsort(a['key'], kind='stable')])
    b = rec16(200000, 82, 1000)
    note("noalloc direct n=%d P=16" % len(b))
    out = _same(b, even(len(b), 16), [0] + even(len(b), 15), 'key')
    assert numpy.array_equal(out, b[numpy.argsort(b['key'], kind='stable')])


def test_noalloc_packed_and_inplace_two_calls():
    """packed records (24 bytes, signed key) out of place and in place"""
    g = numpy.random.default_rng(83)
    for n, P in ((500000, 8), (70000, 3)):
        a = numpy.zeros(n, dtype=[('key', 'i8'), ('a', 'i8'), ('b', 'i8')])
        a['key'] = g.integers(-500, 500, n)
        a['a'] = numpy.arange(n)
        note("noalloc packed n=%d P=%d" % (n, P))
        _same(a, even(n, P), even(n, P), 'key')
        out = _same(a, even(n, P), even(n, P), 'key', inplace=True)
        assert numpy.array_equal(out, a[numpy.argsort(a['key'], kind='stable')])


def test_noalloc_one_rank_and_empty():
    a = rec16(100000, 84, 1 << 20)
    _same(a, [len(a)], [len(a)], 'key')
    e = numpy.zeros(0, dtype=a.dtype)
    z = _counts([0, 0])
    k = key_desc(a.dtype, 'key')
    _lib.check(lib.mpsort_gpu_newarray_noalloc(None, z, 2, None, z, 16, ctypes.byref(k), 0))


def test_noalloc_rejects_a_sort_larger_than_its_scratch():
    """the limit is checked before any device work: a sort needing more than
    the scratch array returns MPSORT_ERR_NOMEM with the need in the message
    (no device buffer of that size is touched)"""
    k = key_desc(numpy.dtype([('key', 'u8'), ('idx', 'u8')]), 'key')
    n = 1 << 24
    buf = lib.mpsort_gpu_dmalloc(64)
    try:
        rc = lib.mpsort_gpu_newarray_noalloc(buf, _counts([n]), 1, buf, _counts([n]), 16, ctypes.byref(k), 0)
    finally:
        lib.mpsort_gpu_dfree(buf)
    assert rc == _lib.MPSORT_ERR_NOMEM
    assert "workspace" in lib.mpsort_gpu_last_error().decode()
