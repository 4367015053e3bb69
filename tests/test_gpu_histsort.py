"""GPU parity of the histogram sort against the oracle (C restatement pinned by
tests/test_oracle.py) and against the compiled reference: output bytes and
every per-virtual-rank count, bit-exact."""
import sys

import numpy
import pytest

import oracle
from gpu_util import even, gpu_sort, gpu_sort_device, last_counts
from mpsort import workloads

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rec16(n, seed, nkeys=None):
    g = numpy.random.default_rng(seed)
    a = numpy.zeros(n, workloads.REC16)
    if nkeys:
        a['key'] = g.integers(0, nkeys, n)
    else:
        a['key'] = g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
    a['idx'] = numpy.arange(n)
    return a


def check_parity(a, nin, nout, field, flags=0, device=False):
    note("case n=%d P=%d itemsize=%d field=%s flags=%d device=%s" % (len(a), len(nin), a.dtype.itemsize,
                                                                     field, flags, device))
    key = oracle.key_for(a.dtype, field)
    ref, info = oracle.histogram_sort(a, nin, nout, key)
    assert info['rc'] == 0
    if device:
        out, st = gpu_sort_device(a, nin, nout, field, flags)
    else:
        out, st = gpu_sort(a, nin, nout, field, flags)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8)), "record bytes differ"
    P = len(nin)
    if P > 1 and len(a):
        c = last_counts(P, 8 * key.nwords)
        assert numpy.array_equal(c['C'], info['C'])
        assert numpy.array_equal(c['P'], info['P']), "splitters differ"
        assert numpy.array_equal(c['Pmin'], info['Pmin'])
        assert numpy.array_equal(c['Pmax'], info['Pmax'])
        assert numpy.array_equal(c['myCLT'], info['myCLT'])
        assert numpy.array_equal(c['myCLE'], info['myCLE'])
        assert numpy.array_equal(c['myC'], info['myC'])
        assert numpy.array_equal(c['send'], info['send'])
    if not (flags & 2):
        assert st.checksum_in == info['checksum_in'] == st.checksum_out
    return out, st


def test_config1_shape_small():
    """4 virtual ranks of {u64 key, u64 index}, PCG64 keys (the 4 x 1e6
    workload, cut to 4 x 2000)."""
    n = 8000
    a = workloads.records16(workloads.pcg64(n))
    out, st = check_parity(a, even(n, 4), even(n, 4), 'key')
    assert st.first_passes == 8


@pytest.mark.parametrize("P", [1, 2, 3, 4, 7, 16, 64, 65, 130, 300])
def test_rank_counts(P):
    n = 5000 + 37 * P
    a = rec16(n, P)
    g = numpy.random.default_rng(100 + P)
    cuts = numpy.sort(g.integers(0, n + 1, P - 1))
    nout = numpy.diff(numpy.concatenate([[0], cuts, [n]]))
    cuts2 = numpy.sort(g.integers(0, n + 1, P - 1))
    nin = numpy.diff(numpy.concatenate([[0], cuts2, [n]]))
    check_parity(a, nin, nout, 'key')


@pytest.mark.parametrize("nkeys", [1, 2, 7, 1000])
def test_duplicates(nkeys):
    n = 20000
    a = rec16(n, 7 + nkeys, nkeys)
    check_parity(a, even(n, 8), even(n, 8), 'key')


@pytest.mark.parametrize("P", [70, 200])
def test_duplicates_many_ranks(P):
    """Few key values over more ranks than one wave has lanes: the layout's
    fill from the equal-key supplies crosses the 64-sender chunks."""
    n = 30000
    a = rec16(n, 11 + P, 3)
    g = numpy.random.default_rng(7 * P)
    nin = numpy.diff(numpy.concatenate([[0], numpy.sort(g.integers(0, n + 1, P - 1)), [n]]))
    nout = numpy.diff(numpy.concatenate([[0], numpy.sort(g.integers(0, n + 1, P - 1)), [n]]))
    check_parity(a, nin, nout, 'key')


def test_empty_ranks_and_outputs():
    a = rec16(1000, 3, 100)
    check_parity(a, [0, 400, 0, 600], [200, 200, 0, 600], 'key')
    check_parity(a, [1000, 0, 0, 0], [0, 0, 0, 1000], 'key')


@pytest.mark.parametrize("dt", ['i4', 'u4', 'i8', 'u8'])
def test_scalar_arrays(dt):
    g = numpy.random.default_rng(5)
    info = numpy.iinfo(dt)
    a = g.integers(info.min, info.max, 30000, dtype=dt, endpoint=True)
    check_parity(a, even(len(a), 5), even(len(a), 5), None)


def test_key_at_offset_and_signed_struct():
    dt = numpy.dtype([('value', 'i8'), ('key', 'i8')])
    g = numpy.random.default_rng(1234)
    a = numpy.zeros(10000, dt)
    a['value'] = numpy.arange(10000)
    a['key'] = g.integers(-400, 600, 10000)
    check_parity(a, even(10000, 4), even(10000, 4), 'key')


def test_32_byte_particles():
    n = 12000
    a = numpy.zeros(n, workloads.PARTICLE32)
    a['id'] = workloads.permutation(n, 9)
    a['pos'] = numpy.random.default_rng(1).random((n, 3), dtype=numpy.float32)
    check_parity(a, even(n, 6), even(n, 6), 'id')


def test_general_path_few_items():
    """36-byte records with a 3-word u8 key (test_mpsort.py:330-351 shape)."""
    dt = numpy.dtype([('vkey', ('u8', 3)), ('vector', ('u4', 3))])
    g = numpy.random.default_rng(11)
    a = numpy.zeros(3000, dt)
    a['vkey'] = g.integers(0, 4, (3000, 3))
    a['vector'] = g.integers(0, 100, (3000, 3))
    out, st = check_parity(a, [0, 1000, 2000, 0], [1500, 0, 0, 1500], 'vkey')
    assert st.general_path == 1


def test_general_path_issue7_shape():
    """40-byte records with a 2-word radix (test_mpsort.py:531-556); the
    fixture there is a pickle, so synthetic rows of the same shape are used:
    12 ranks with row counts [0,0,2,7,1,0,4,0,0,6,14,5]."""
    counts = [0, 0, 2, 7, 1, 0, 4, 0, 0, 6, 14, 5]
    n = sum(counts)
    g = numpy.random.default_rng(7)
    A = g.integers(0, 2 ** 32, (n, 10), dtype=numpy.uint64).astype(numpy.uint32)
    A[:, 0] %= 3  # most significant word: many ties
    s = numpy.zeros(n, dtype=[('radix', ('u8', 2)), ('ext', ('u8', 3))])
    s['radix'][:, 0] = A[:, 5].astype('u8') + (A[:, 4].astype('u8') << numpy.uint64(32))
    s['radix'][:, 1] = A[:, 0]
    s['ext'] = g.integers(0, 2 ** 63, (n, 3))
    out, st = check_parity(s, counts, counts, 'radix')
    S = s[numpy.lexsort(s['radix'].T)]
    assert numpy.array_equal(out['radix'], S['radix'])


def test_forced_general_matches_direct():
    a = rec16(20000, 21, 5000)
    check_parity(a, even(20000, 4), even(20000, 4), 'key', flags=16)


def test_device_pointers_and_inplace():
    a = rec16(50000, 31)
    check_parity(a, even(50000, 8), even(50000, 8), 'key', device=True)
    key = oracle.key_for(a.dtype, 'key')
    ref, _ = oracle.histogram_sort(a, even(50000, 8), even(50000, 8), key)
    b = a.copy()
    gpu_sort(b, even(50000, 8), even(50000, 8), 'key', inplace=True)
    assert numpy.array_equal(b, ref)


def test_radix_callback_form():
    import ctypes
    from mpsort._lib import RADIX_FN, lib, check
    a = rec16(6000, 41, 300)

    def radix(ptr, out, arg):
        ctypes.memmove(out, ptr, 8)

    cb = RADIX_FN(radix)
    nin = (ctypes.c_int64 * 3)(*even(6000, 3))
    out = numpy.zeros_like(a)
    check(lib.mpsort_gpu_newarray_radix(a.ctypes.data, nin, 3, out.ctypes.data, nin, 16, cb, 8, None, 0))
    ref = a[numpy.argsort(a['key'], kind='stable')]
    assert numpy.array_equal(out, ref)


def test_repeat_is_deterministic():
    a = rec16(200000, 51, 1 << 12)
    o1, _ = gpu_sort(a, even(len(a), 16), even(len(a), 16), 'key')
    o2, _ = gpu_sort(a, even(len(a), 16), even(len(a), 16), 'key')
    assert numpy.array_equal(o1, o2)


def test_config1_full_against_reference():
    """4 x 1,000,000 {u64 key, u64 index} PCG64 records: bytes equal the
    compiled reference's mpsort_omp output and the oracle's."""
    n = 4000000
    a = workloads.records16(workloads.pcg64(n))
    note("config1 full n=%d" % n)
    out, st = gpu_sort(a, even(n, 4), even(n, 4), 'key')
    key = oracle.key_for(a.dtype, 'key')
    if oracle.ref_available():
        ref = oracle.ref_mpsort_omp(a, key, nthreads=4)
    else:
        ref, _ = oracle.histogram_sort(a, even(n, 4), even(n, 4), key)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))


def test_clustered_large_properties():
    """2^26 clustered Morton records on 64 virtual ranks: sorted, a
    permutation of the input, stable (ties in input order)."""
    n = 1 << 26
    a = workloads.clustered_records(n, seed=2)
    note("clustered n=%d" % n)
    out, st = gpu_sort_device(a, even(n, 64), even(n, 64), 'key')
    k = out['key']
    assert bool(numpy.all(k[1:] >= k[:-1]))
    ties = k[1:] == k[:-1]
    assert bool(numpy.all(out['idx'][1:][ties] > out['idx'][:-1][ties]))
    assert numpy.array_equal(numpy.sort(out['idx']), numpy.arange(n))
    assert st.checksum_in == st.checksum_out


def _device_inplace(a, nin, field, flags=0):
    """mpsort_gpu_newarray with base == out in device memory"""
    import ctypes
    from gpu_util import key_desc
    from mpsort import _lib
    from mpsort._lib import lib
    nb = a.nbytes
    d = lib.mpsort_gpu_dmalloc(max(nb, 1))
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, nb))
        P = len(nin)
        c = (ctypes.c_int64 * P)(*[int(x) for x in nin])
        k = key_desc(a.dtype, field)
        _lib.check(lib.mpsort_gpu_newarray(d, c, P, d, c, a.dtype.itemsize, ctypes.byref(k),
                                           flags | _lib.DEVICE_PTRS))
        out = numpy.empty_like(a)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, d, nb))
    finally:
        lib.mpsort_gpu_dfree(d)
    return out, _lib.last_stats()


def test_device_inplace_packed_24_byte_records():
    """In-place sort of device memory on the packed-record path (24-byte
    records): the gather reads the caller's records from a copy, so writing
    the output over the input cannot corrupt it."""
    g = numpy.random.default_rng(61)
    n = 300000
    a = numpy.zeros(n, dtype=[('key', 'i8'), ('a', 'i8'), ('b', 'i8')])
    a['key'] = g.integers(-1000, 1000, n)
    a['a'] = numpy.arange(n)
    a['b'] = g.integers(0, 2 ** 62, n)
    note("device in-place packed n=%d itemsize=24" % n)
    out, st = _device_inplace(a, even(n, 8), 'key')
    assert st.general_path == 1
    assert numpy.array_equal(out, a[numpy.argsort(a['key'], kind='stable')])
    assert st.checksum_in == st.checksum_out


def test_device_inplace_forced_general():
    a = rec16(300000, 62, 1 << 10)
    note("device in-place FORCE_GENERAL n=%d" % len(a))
    out, st = _device_inplace(a, even(len(a), 8), 'key', flags=16)
    assert st.general_path == 1
    assert numpy.array_equal(out, a[numpy.argsort(a['key'], kind='stable')])


def _callback_sort(a, nin, nout, offset, nwords, flags=0):
    import ctypes
    from mpsort import _lib, workloads as w
    from mpsort._lib import lib
    gen = w._gen()
    fn = ctypes.cast(gen.mpsort_radix_u64_words, _lib.RADIX_FN)
    arg = (ctypes.c_int64 * 2)(offset, nwords)
    P = len(nin)
    ci = (ctypes.c_int64 * P)(*[int(x) for x in nin])
    co = (ctypes.c_int64 * P)(*[int(x) for x in nout])
    out = numpy.zeros_like(a)
    _lib.check(lib.mpsort_gpu_newarray_radix(a.ctypes.data, ci, P, out.ctypes.data, co, a.dtype.itemsize, fn,
                                             8 * nwords, ctypes.cast(arg, ctypes.c_void_p), flags))
    return out, _lib.last_stats()


def test_radix_callback_at_scale():
    """The opaque-callback path (radix materialised by host threads) at 2^22
    records over 16 ranks with a C callback, against the oracle's bytes and
    the per-rank counts"""
    n, P = 1 << 22, 16
    a = rec16(n, 63, 1 << 20)
    note("callback n=%d P=%d" % (n, P))
    out, st = _callback_sort(a, even(n, P), even(n, P), 0, 1)
    key = oracle.key_for(a.dtype, 'key')
    ref, info = oracle.histogram_sort(a, even(n, P), even(n, P), key)
    assert numpy.array_equal(out, ref)
    c = last_counts(P, 8)
    assert numpy.array_equal(c['P'], info['P'])
    assert numpy.array_equal(c['send'], info['send'])
    assert st.checksum_in == info['checksum_in'] == st.checksum_out


def test_radix_callback_three_word_radix():
    """a 24-byte radix (3 u64 words, the last most significant) of 32-byte
    records, uneven partitions"""
    g = numpy.random.default_rng(64)
    n = 200000
    a = numpy.zeros(n, dtype=[('k', ('u8', 3)), ('v', 'u8')])
    a['k'][:, 0] = g.integers(0, 2 ** 63, n)
    a['k'][:, 1] = g.integers(0, 3, n)
    a['k'][:, 2] = g.integers(0, 5, n)
    a['v'] = numpy.arange(n)
    nin = [0, n // 3, n - n // 3, 0]
    nout = [n // 2, 0, n - n // 2 - 7, 7]
    note("callback 3-word radix n=%d" % n)
    out, st = _callback_sort(a, nin, nout, 0, 3)
    assert numpy.array_equal(out, a[numpy.lexsort(a['k'].T)])


@pytest.mark.parametrize("entry", ["radix_sort", "mpsort_omp"])
def test_radix_sort_and_mpsort_omp_exports(entry):
    """The reference's single-process entry points (mpsort.h:1-11) with a C
    radix callback: 16-byte records sorted in place by a u64 radix, equal to
    numpy's stable sort (ties keep their input order)."""
    import ctypes
    from mpsort import _lib, workloads as w
    from mpsort._lib import lib
    n = 300000
    a = rec16(n, 71, 1 << 12)
    want = a[numpy.argsort(a['key'], kind='stable')]
    fn = ctypes.cast(w._gen().mpsort_radix_u64_words, _lib.RADIX_FN)
    arg = (ctypes.c_int64 * 2)(0, 1)
    f = getattr(lib, entry)
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, _lib.RADIX_FN, ctypes.c_size_t,
                  ctypes.c_void_p]
    note("%s n=%d" % (entry, n))
    f(a.ctypes.data, n, a.dtype.itemsize, fn, 8, ctypes.cast(arg, ctypes.c_void_p))
    assert numpy.array_equal(a, want)


@pytest.mark.parametrize("nwords", [5, 8])
def test_radix_callback_long_radix(nwords):
    """radixes of 40 and 64 bytes (5 and 8 u64 words, the last most
    significant), longer than one packed sort record: sorted in 24-byte
    chunks, equal to numpy's lexsort (ties in input order) on uneven
    partitions"""
    g = numpy.random.default_rng(65 + nwords)
    n = 150000
    a = numpy.zeros(n, dtype=[('k', ('u8', nwords)), ('v', 'u8')])
    for w in range(nwords):
        a['k'][:, w] = g.integers(0, 3 if w >= nwords - 3 else 2 ** 64, n, dtype=numpy.uint64)
    a['v'] = numpy.arange(n)
    nin = [n // 4, 0, n - n // 4]
    nout = [n // 3, n - n // 3 - 11, 11]
    note("callback %d-byte radix n=%d" % (8 * nwords, n))
    out, st = _callback_sort(a, nin, nout, 0, nwords)
    order = numpy.lexsort([a['v']] + [a['k'][:, w] for w in range(nwords)])
    assert numpy.array_equal(out, a[order])
    assert st.checksum_in == st.checksum_out
