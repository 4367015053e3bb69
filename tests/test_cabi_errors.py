"""Argument validation of the C ABI and of the Python binding, on the CPU.

Every entry point checks its arguments before it touches a device, so these
run without a GPU: each bad call must return its MPSORT_ERR_* code with a
message (the reference aborts on the same conditions instead,
mpsort-mpi.c:74,185,225-232,329), and the Python binding must raise the
exception type binding.pyx raises for it (binding.pyx:65,79,157-189).  A
well-formed call on a machine without a GPU fails loudly with
MPSORT_ERR_NODEVICE: there is no CPU fallback."""
import ctypes

import numpy
import pytest

import mpsort
from mpsort import _lib
from mpsort._lib import KeyDesc, lib

I64 = ctypes.c_int64


def counts(*v):
    return (I64 * len(v))(*v)


def key(offset=0, width=8, signed=0, nwords=1):
    return KeyDesc(offset=offset, width=width, is_signed=signed, nwords=nwords, reserved=0)


def newarray(nin, nout, elsize=16, k=None, flags=0, base=None, out=None, P=None):
    buf = numpy.zeros(max(1, sum(nin)) * elsize, numpy.uint8)
    b = buf.ctypes.data if base is None else base
    o = buf.ctypes.data if out is None else out
    kp = ctypes.byref(k if k is not None else key()) if k is not False else None
    return lib.mpsort_gpu_newarray(b, counts(*nin), len(nin) if P is None else P, o, counts(*nout), elsize, kp,
                                   flags)


def err():
    return lib.mpsort_gpu_last_error().decode()


@pytest.mark.parametrize("case, code, text", [
    (dict(nin=[4], nout=[4], P=0), _lib.MPSORT_ERR_ARG, "virtual ranks"),
    (dict(nin=[4], nout=[4], P=5000), _lib.MPSORT_ERR_ARG, "virtual ranks"),
    (dict(nin=[4, -1], nout=[2, 1]), _lib.MPSORT_ERR_ARG, "negative count"),
    (dict(nin=[4, 4], nout=[4, 5]), _lib.MPSORT_ERR_SIZE_MISMATCH, "does not match"),
    (dict(nin=[4], nout=[4], k=False), _lib.MPSORT_ERR_ARG, "NULL"),
    (dict(nin=[4], nout=[4], elsize=0), _lib.MPSORT_ERR_ARG, "elsize"),
    (dict(nin=[4], nout=[4], k=key(width=2)), _lib.MPSORT_ERR_UNSUPPORTED, "width"),
    (dict(nin=[4], nout=[4], k=key(nwords=0)), _lib.MPSORT_ERR_ARG, "nwords"),
    (dict(nin=[4], nout=[4], k=key(offset=12)), _lib.MPSORT_ERR_ARG, "outside"),
    (dict(nin=[4], nout=[4], k=key(offset=-4)), _lib.MPSORT_ERR_ARG, "outside"),
    (dict(nin=[4], nout=[4], elsize=40, k=key(nwords=5)), _lib.MPSORT_ERR_UNSUPPORTED, "exceeds 32"),
    (dict(nin=[4], nout=[4], elsize=40, k=key(nwords=4)), _lib.MPSORT_ERR_UNSUPPORTED, "packed record"),
    (dict(nin=[4], nout=[4], base=0), _lib.MPSORT_ERR_ARG, "NULL data"),
])
def test_newarray_argument_errors(case, code, text):
    assert newarray(**case) == code
    assert text in err()


def test_size_mismatch_message_names_both_totals():
    assert newarray([3, 4], [5, 5]) == _lib.MPSORT_ERR_SIZE_MISMATCH
    assert "10 != 7" in err()


def test_python_check_maps_codes_to_the_binding_exceptions():
    assert newarray([4, 4], [4, 5]) == _lib.MPSORT_ERR_SIZE_MISMATCH
    with pytest.raises(ValueError):
        _lib.check(_lib.MPSORT_ERR_SIZE_MISMATCH)
    with pytest.raises(ValueError):
        _lib.check(_lib.MPSORT_ERR_ARG)
    with pytest.raises(TypeError):
        _lib.check(_lib.MPSORT_ERR_UNSUPPORTED)
    with pytest.raises(_lib.MPSortError):
        _lib.check(_lib.MPSORT_ERR_CHECKSUM)


RADIX = _lib.RADIX_FN(lambda p, r, a: None)


@pytest.mark.parametrize("rsize, flags, fn, code", [
    (8, 0, None, _lib.MPSORT_ERR_ARG),
    (0, 0, RADIX, _lib.MPSORT_ERR_UNSUPPORTED),
    (8, _lib.DEVICE_PTRS, RADIX, _lib.MPSORT_ERR_ARG),
])
def test_radix_callback_argument_errors(rsize, flags, fn, code):
    buf = numpy.zeros(64, numpy.uint8)
    f = fn if fn is not None else ctypes.cast(None, _lib.RADIX_FN)
    rc = lib.mpsort_gpu_newarray_radix(buf.ctypes.data, counts(4), 1, buf.ctypes.data, counts(4), 16, f, rsize,
                                       None, flags)
    assert rc == code, err()


def test_long_radix_is_accepted():
    """a radix longer than the packed record (radixsort.c takes any rsize):
    sorted in 24-byte chunks, so not rejected (no GPU here: NODEVICE)"""
    buf = numpy.zeros(64, numpy.uint8)
    rc = lib.mpsort_gpu_newarray_radix(buf.ctypes.data, counts(4), 1, buf.ctypes.data, counts(4), 16, RADIX, 40,
                                       None, 0)
    assert rc != _lib.MPSORT_ERR_UNSUPPORTED
    if _lib.device_count() == 0:
        assert rc == _lib.MPSORT_ERR_NODEVICE


def test_learned_entry_rejects_a_null_model():
    buf = numpy.zeros(64, numpy.uint8)
    rc = lib.mpsort_gpu_newarray_learned(buf.ctypes.data, counts(4), 1, buf.ctypes.data, counts(4), 16,
                                         ctypes.byref(key()), None, 0)
    assert rc == _lib.MPSORT_ERR_ARG and "model" in err()


def test_no_gpu_means_a_loud_error():
    """A well-formed call without a GPU returns MPSORT_ERR_NODEVICE: nothing
    sorts on the CPU in the product library."""
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    a = numpy.arange(8, dtype=numpy.uint64)[::-1].copy()
    rc = lib.mpsort_gpu_newarray(a.ctypes.data, counts(8), 1, a.ctypes.data, counts(8), 8, ctypes.byref(key()), 0)
    assert rc == _lib.MPSORT_ERR_NODEVICE
    with pytest.raises(_lib.MPSortError):
        mpsort.sort(a)


# ---- the Python binding's own checks (binding.pyx:157-189), before any call
def test_binding_rejects_non_contiguous_data():
    a = numpy.zeros(16, dtype=[('key', 'i8'), ('v', 'i8')])[::2]
    with pytest.raises(ValueError, match="C_CONTIGUOUS"):
        mpsort.binding.sort(a, 'key')


def test_binding_rejects_non_contiguous_out():
    a = numpy.zeros(8, dtype=[('key', 'i8'), ('v', 'i8')])
    out = numpy.zeros(16, dtype=a.dtype)[::2]
    with pytest.raises(ValueError, match="C_CONTIGUOUS"):
        mpsort.binding.sort(a, 'key', out=out)


def test_binding_rejects_changed_total_size():
    a = numpy.zeros(8, dtype=[('key', 'i8'), ('v', 'i8')])
    with pytest.raises(ValueError, match="total size of array changed"):
        mpsort.binding.sort(a, 'key', out=numpy.zeros(7, dtype=a.dtype))


def test_binding_rejects_item_size_mismatch():
    a = numpy.zeros(8, dtype=[('key', 'i8'), ('v', 'i8')])
    with pytest.raises(ValueError, match="item size mismatch"):
        mpsort.binding.sort(a, 'key', out=numpy.zeros(8, dtype='i8'))


@pytest.mark.parametrize("dt", ['f8', 'i2', 'u1', 'f4'])
def test_binding_rejects_non_integer_or_short_keys(dt):
    a = numpy.zeros(8, dtype=[('key', dt), ('v', 'i8')])
    with pytest.raises(TypeError):
        mpsort.binding.sort(a, 'key')


def test_binding_rejects_2d_subarray_keys():
    a = numpy.zeros(8, dtype=[('key', 'i8', (2, 2)), ('v', 'i8')])
    with pytest.raises(ValueError, match="not 1d"):
        mpsort.binding.sort(a, 'key')


def test_binding_rejects_unknown_field():
    a = numpy.zeros(8, dtype=[('key', 'i8'), ('v', 'i8')])
    with pytest.raises((KeyError, ValueError)):
        mpsort.binding.sort(a, 'nope')


def test_binding_rejects_foreign_communicators():
    a = numpy.zeros(8, dtype=[('key', 'i8'), ('v', 'i8')])
    with pytest.raises(ValueError, match="VirtualComm"):
        mpsort.binding.sort(a, 'key', comm=object())


def test_virtual_comm_partition_checks():
    with pytest.raises(ValueError):
        mpsort.VirtualComm(vranks=0)
    with pytest.raises(ValueError):
        mpsort.VirtualComm(in_sizes=[1, 2], out_sizes=[3])
    c = mpsort.VirtualComm(in_sizes=[1, 2])
    with pytest.raises(ValueError, match="do not add up"):
        c.partition(4, c.in_sizes)
    assert mpsort.VirtualComm(vranks=3).partition(10, None) == [3, 3, 4]


def test_empty_sort_returns_before_any_device_call():
    """test_mpsort.py:319-327: an empty array with a 3-word key sorts (on any
    machine: nothing is sent to the device)."""
    s = numpy.empty(0, dtype=[('vkey', ('u8', 3)), ('vector', ('u4', 3))])
    for tuning in ([], ['DISABLE_SPARSE_ALLTOALLV'], ['REQUIRE_SPARSE_ALLTOALLV'], ['REQUIRE_GATHER_SORT'],
                   ['DISABLE_GATHER_SORT']):
        mpsort.sort(s, 'vkey', out=s, comm=mpsort.VirtualComm(vranks=4), tuning=tuning)


def test_histogram_of_nothing():
    """test_mpsort.py:312-316"""
    h = mpsort.histogram([], [1], mpsort.VirtualComm())
    assert list(h) == [0, 0]


def test_histogram_follows_digitize():
    g = numpy.random.default_rng(5)
    x = g.integers(-20, 20, 500)
    edges = numpy.array([-10, -3, 0, 4, 11])
    for right in (False, True):
        h = mpsort.histogram(x, edges, mpsort.VirtualComm(), right=right)
        want = numpy.zeros(len(edges) + 1, int)
        for v in x:
            # brute force: the number of edges below (or at, right=False) v
            want[sum(1 for e in edges if (e <= v if not right else e < v))] += 1
        assert list(h) == list(want)


def test_globalrange_and_globalindices_on_one_process():
    comm = mpsort.VirtualComm(vranks=4)
    a = numpy.zeros(10)
    assert mpsort.globalrange(a, comm) == (0, 10)
    gi = mpsort.globalindices(a, comm)
    assert gi.dtype == numpy.int32 and list(gi) == list(range(10))


def test_version():
    """test_mpsort.py:309-311"""
    assert hasattr(mpsort, "__version__")


def test_fault_hook_arguments():
    """mpsort_gpu_set_faults checks its bits and the capacity; 0 clears"""
    assert lib.mpsort_gpu_set_faults(1 << 9, 0) == _lib.MPSORT_ERR_ARG and "unknown fault" in err()
    assert lib.mpsort_gpu_set_faults(_lib.FAULT_BUCKET_CAP, 0) == _lib.MPSORT_ERR_ARG and "bucket_cap" in err()
    assert lib.mpsort_gpu_set_faults(_lib.FAULT_BUCKET_CAP, 9217) == _lib.MPSORT_ERR_ARG
    assert lib.mpsort_gpu_set_faults(_lib.FAULT_BUCKET_CAP | _lib.FAULT_PLAIN, 100) == 0
    assert lib.mpsort_gpu_set_faults(0, 0) == 0
