"""The plain path's SecondSort without an exchange (at most 64 ranks, keys of
at most 8 bytes, 8- or 16-byte sort records): receiver tables from an even
sample of every receiver's records, each receiver bucket gathered from the
senders' sorted segments in sender order and sorted in LDS.  The exchange
followed by a stable sort of each receiver's records (mpsort-mpi.c:571-597)
fixes the result; these tests hold the merged form to the oracle's bytes,
counts and checksums, and to the explicit exchange's output, on inputs that
strain the sampled tables: heavy keys, few values, sorted and reversed
input, tiny and empty receivers, and the forced fallbacks."""
import sys

import numpy
import pytest

import oracle
from gpu_util import even, gpu_sort
from mpsort import _lib, workloads
from mpsort._lib import lib
from test_gpu_histsort import check_parity

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def no_faults_or_options():
    # recorded exchange options would ask for an exchange
    lib.mpsort_mpi_unset_options(-1)
    yield
    _lib.check(lib.mpsort_gpu_set_faults(0, 0))


def keys(shape, n, seed):
    g = numpy.random.default_rng(seed)
    if shape == "random":
        return g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
    if shape == "heavy":
        # half the records share one key, the rest spread out
        k = g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
        k[g.random(n) < 0.5] = numpy.uint64(0x8000000000000000)
        return k
    if shape == "few":
        return g.integers(0, 5, n).astype(numpy.uint64) * numpy.uint64(3 << 40)
    if shape == "zero_heavy":
        k = g.integers(0, 2 ** 20, n, dtype=numpy.uint64)
        k[g.random(n) < 0.3] = 0
        return k
    if shape == "sorted":
        return numpy.sort(g.integers(0, 2 ** 64, n, dtype=numpy.uint64))
    if shape == "reversed":
        return numpy.sort(g.integers(0, 2 ** 64, n, dtype=numpy.uint64))[::-1].copy()
    if shape == "narrow":
        # keys within a range of 2^12 values: many repeats, no one heavy key
        return g.integers(2 ** 40, 2 ** 40 + 4096, n, dtype=numpy.uint64)
    if shape == "two_scales":
        # half the senders' keys tiny, the other half huge (every receiver
        # near the middle draws from both)
        k = g.integers(0, 2 ** 16, n, dtype=numpy.uint64)
        k[n // 2:] = g.integers(2 ** 63, 2 ** 64, n - n // 2, dtype=numpy.uint64)
        return k
    raise ValueError(shape)


SHAPES = ["random", "heavy", "few", "zero_heavy", "sorted", "reversed", "narrow", "two_scales"]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("P", [2, 5, 64])
def test_merged_matches_the_oracle_and_the_exchange(shape, P):
    n = 60000 + 101 * P
    a = workloads.records16(keys(shape, n, 7 + P))
    note("plain merged shape=%s P=%d n=%d" % (shape, P, n))
    out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    assert st.second_merged == 1 and st.exchange_runs == 0 and st.second_passes == 0
    out2, st2 = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.EXCHANGE, device=True)
    assert st2.second_merged == 0
    assert numpy.array_equal(out, out2)


@pytest.mark.parametrize("P", [3, 64])
def test_ragged_and_empty_receivers(P):
    n = 40000
    g = numpy.random.default_rng(P)
    a = workloads.records16(g.integers(0, 2 ** 64, n, dtype=numpy.uint64))
    nin = numpy.diff(numpy.concatenate([[0], numpy.sort(g.integers(0, n + 1, P - 1)), [n]]))
    nout = numpy.zeros(P, numpy.int64)
    nout[0], nout[-1] = n // 3, n - n // 3  # every rank between receives nothing
    out, st = check_parity(a, nin, nout, 'key', device=True)
    assert st.second_merged == 1


@pytest.mark.parametrize("kind", ["u8", "i8", "u4_pairs"])
def test_eight_byte_records(kind):
    """scalar 8-byte records and 4-byte-key pairs: the 8-byte record forms"""
    n, P = 50000, 16
    g = numpy.random.default_rng(11)
    if kind == "u8":
        a = g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
        field = None
    elif kind == "i8":
        a = g.integers(-2 ** 63, 2 ** 63, n, dtype=numpy.int64)
        field = None
    else:
        a = numpy.zeros(n, [('key', 'u4'), ('v', 'u4')])
        a['key'] = g.integers(0, 1000, n)
        a['v'] = numpy.arange(n)
        field = 'key'
    note("plain merged 8-byte records kind=%s" % kind)
    key = oracle.key_for(a.dtype, field)
    ref, info = oracle.histogram_sort(a, even(n, P), even(n, P), key)
    out, st = gpu_sort(a, even(n, P), even(n, P), field)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))
    assert st.second_merged == 1
    assert st.checksum_in == info['checksum_in'] == st.checksum_out


def test_general_path_records():
    """24-byte records: packed {key, index} sort records take the merged form"""
    n, P = 30000, 8
    g = numpy.random.default_rng(5)
    a = numpy.zeros(n, [('x', 'f8'), ('key', 'i8'), ('w', 'u8')])
    a['key'] = g.integers(-50, 50, n)
    a['x'] = g.random(n)
    a['w'] = numpy.arange(n)
    out, st = check_parity(a, even(n, P), even(n, P), 'key')
    assert st.general_path == 1 and st.second_merged == 1


def test_large_receivers():
    """2^22 clustered records on 64 ranks: receivers of 65536 records, more
    than the 12288 samples of their tables"""
    n, P = 1 << 22, 64
    a = workloads.clustered_records(n, seed=4)
    note("plain merged clustered n=%d" % n)
    out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    assert st.second_merged == 1 and st.plain_sides == 0


@pytest.mark.parametrize("cap, global_fallback", [(32, False), (1, True)])
def test_forced_bucket_overflow(cap, global_fallback):
    """receiver buckets above a forced capacity (they hold 21 records on
    average): the bucket fallback for some hundreds of them, or with most
    buckets oversized the radix sort of every receiver's whole segment"""
    n, P = 1 << 19, 8
    a = workloads.records16(keys("random", n, 3))
    _lib.check(lib.mpsort_gpu_set_faults(_lib.FAULT_BUCKET_CAP, cap))
    note("plain merged bucket cap %d" % cap)
    out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    assert st.second_merged == 1 and st.buckets_overflow > 0
    assert st.plain_sides == (1 if global_fallback else 0)
    assert global_fallback or st.buckets_overflow <= 4096


def test_forced_plain_fallback():
    n, P = 1 << 18, 16
    a = workloads.records16(keys("heavy", n, 9))
    _lib.check(lib.mpsort_gpu_set_faults(_lib.FAULT_PLAIN, 0))
    out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    assert st.second_merged == 1 and st.plain_sides == 1


# The plain FirstSort as a bucket sort: 1024 buckets per rank from a sorted
# sample of 65536 of its keys, then the in-LDS bucket sorts (mpsort-mpi.c:369
# asks only for a stable local sort).  FIRST_BUCKETS takes it below its size
# threshold, FIRST_RADIX keeps the digit passes.

def test_bucket_first_sort_back_to_back_shapes(monkeypatch):
    """The sample sort's segment arrays stay on the device from one call to
    the next; a call with another rank count or tile shape lays them out
    again."""
    steps = [(64, None), (4, None), (64, None), (7, "512:16"), (7, None), (64, "512:16"), (64, None)]
    for i, (P, tile) in enumerate(steps):
        if tile:
            monkeypatch.setenv("MPSORT_GPU_TILE", tile)
        else:
            monkeypatch.delenv("MPSORT_GPU_TILE", raising=False)
        n = 70000 + 313 * i
        a = workloads.records16(keys("random", n, 101 + i))
        note("plain bucket FirstSort back to back: step %d P=%d tile=%s n=%d" % (i, P, tile, n))
        out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.FIRST_BUCKETS, device=True)
        assert st.first_passes == 0


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("P", [1, 4, 64])
def test_bucket_first_sort_matches_the_oracle(shape, P):
    n = 90000 + 77 * P
    a = workloads.records16(keys(shape, n, 31 + P))
    note("plain bucket FirstSort shape=%s P=%d n=%d" % (shape, P, n))
    out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.FIRST_BUCKETS, device=True)
    assert st.first_passes == 0
    out2, st2 = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.FIRST_RADIX, device=True)
    assert st2.first_passes > 0 or len(numpy.unique(a['key'])) == 1
    assert numpy.array_equal(out, out2)


@pytest.mark.parametrize("kind", ["u8", "i8", "u4_pairs"])
def test_bucket_first_sort_eight_byte_records(kind):
    n, P = 70000, 4
    g = numpy.random.default_rng(13)
    if kind == "u8":
        a = g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
        field = None
    elif kind == "i8":
        a = g.integers(-2 ** 63, 2 ** 63, n, dtype=numpy.int64)
        field = None
    else:
        a = numpy.zeros(n, [('key', 'u4'), ('v', 'u4')])
        a['key'] = g.integers(0, 1000, n)
        a['v'] = numpy.arange(n)
        field = 'key'
    note("plain bucket FirstSort 8-byte records kind=%s" % kind)
    key = oracle.key_for(a.dtype, field)
    ref, info = oracle.histogram_sort(a, even(n, P), even(n, P), key)
    out, st = gpu_sort(a, even(n, P), even(n, P), field, flags=_lib.FIRST_BUCKETS)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))
    assert st.first_passes == 0
    assert st.checksum_in == info['checksum_in'] == st.checksum_out


def test_bucket_first_sort_general_path_inplace():
    """packed records sorted in place (one rank: the FirstSort is the sort)"""
    n = 80000
    g = numpy.random.default_rng(6)
    a = numpy.zeros(n, [('x', 'f8'), ('key', 'i8'), ('w', 'u8')])
    a['key'] = g.integers(-500, 500, n)
    a['w'] = numpy.arange(n)
    for P in (1, 3):
        out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.FIRST_BUCKETS)
        assert st.general_path == 1 and st.first_passes == 0


def test_bucket_first_sort_by_size():
    """2^21 clustered records on 4 ranks (2^19 each): the bucket FirstSort
    without being asked, the merged SecondSort after it"""
    n, P = 1 << 21, 4
    a = workloads.clustered_records(n, seed=6)
    out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    assert st.first_passes == 0 and st.second_merged == 1 and st.plain_sides == 0


@pytest.mark.parametrize("faults, cap", [(_lib.FAULT_BUCKET_CAP, 48), (_lib.FAULT_BUCKET_CAP, 1),
                                         (_lib.FAULT_PLAIN, 0)])
def test_bucket_first_sort_fallbacks(faults, cap):
    """FirstSort buckets above a forced capacity (88 records on average):
    the bucket fallback, or past its limit the radix sort of the whole side;
    and the forced radix sort of both sides"""
    n, P = 1 << 18, 3
    a = workloads.records16(keys("random", n, 17))
    _lib.check(lib.mpsort_gpu_set_faults(faults, cap))
    note("plain bucket FirstSort faults=%d cap=%d" % (faults, cap))
    out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.FIRST_BUCKETS, device=True)
    if faults == _lib.FAULT_PLAIN:
        assert st.plain_sides == 2
    else:
        assert st.buckets_overflow_first > 0
        assert st.plain_sides >= (1 if cap == 1 else 0)


def test_large_32_byte_records_are_packed():
    """32-byte particles in ranks of 2^18 records: sorted as packed {key,
    index} records (bucket FirstSort, merged SecondSort), then gathered"""
    n, P = 1 << 20, 4
    g = numpy.random.default_rng(21)
    a = numpy.zeros(n, [('id', 'u8'), ('pos', 'f4', 3), ('vel', 'f4', 3)])
    a['id'] = g.permutation(n).astype(numpy.uint64)
    a['id'][g.random(n) < 0.01] = 7  # a heavy id as well
    a['pos'] = g.random((n, 3))
    a['vel'] = g.random((n, 3))
    out, st = check_parity(a, even(n, P), even(n, P), 'id', device=True)
    assert st.general_path == 1 and st.first_passes == 0 and st.second_merged == 1
    out2, st2 = check_parity(a, even(n, P), even(n, P), 'id', flags=_lib.FIRST_RADIX, device=True)
    assert st2.general_path == 0 and numpy.array_equal(out, out2)


@pytest.mark.parametrize("P", [1, 4])
@pytest.mark.parametrize("rec", ["rec16", "particle32"])
def test_bucket_first_sort_device_in_place(P, rec):
    """base == out in device memory: the bucket FirstSort reads the input
    before anything writes it (16-byte records), and the packed 32-byte path
    copies the input aside before gathering"""
    from test_gpu_histsort import _device_inplace
    n = 1 << 20 if rec == "particle32" else 120000
    g = numpy.random.default_rng(40 + P)
    if rec == "rec16":
        a = workloads.records16(keys("heavy", n, 40 + P))
        field = 'key'
        flags = _lib.FIRST_BUCKETS
    else:
        a = numpy.zeros(n, workloads.PARTICLE32)
        a['id'] = g.permutation(n).astype(numpy.uint64)
        a['pos'] = g.random((n, 3), dtype=numpy.float32)
        field = 'id'
        flags = 0  # ranks of 2^18 records or more: packed by size
    note("plain bucket FirstSort in place rec=%s P=%d n=%d" % (rec, P, n))
    key = oracle.key_for(a.dtype, field)
    ref, info = oracle.histogram_sort(a, even(n, P), even(n, P), key)
    out, st = _device_inplace(a, even(n, P), field, flags=flags)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))
    assert st.first_passes == 0
    assert st.checksum_in == info['checksum_in'] == st.checksum_out
