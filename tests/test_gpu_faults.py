"""The fault hooks (mpsort_gpu_set_faults) force the paths that the
reference's checks guard: splitters found by the repair rounds, buckets sent
to the overflow fallback, the plain radix sort of a whole side, and data
changed during the sort.  The first three must still give the plain path's
bytes and counts; the last must fail with MPSORT_ERR_CHECKSUM
(mpsort-mpi.c:324-330 aborts there)."""
import sys

import numpy
import pytest

import mpsort
from gpu_util import even, gpu_sort, last_counts
from mpsort import _lib, learned, workloads
from mpsort._lib import lib
from test_gpu_learned import keys_of, trained

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def no_faults_after():
    yield
    _lib.check(lib.mpsort_gpu_set_faults(0, 0))


def learned_vs_plain(faults, cap, n, P, dist="uniform", flags=0):
    model = trained(dist, n)
    a = workloads.records16(keys_of(dist, n, 5))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    ref_counts = last_counts(P, 8)
    _lib.check(lib.mpsort_gpu_set_faults(faults, cap))
    try:
        out = learned.sort_learned(a, model, 'key', nin=even(n, P), flags=flags)
        st = mpsort.last_stats()
        counts = last_counts(P, 8)
    finally:
        model.close()
    assert numpy.array_equal(out, ref), "learned output differs from the plain path"
    for k in ("C", "P", "myCLT", "myCLE", "myC", "send"):
        assert numpy.array_equal(counts[k], ref_counts[k]), k
    assert st['checksum_in'] == st['checksum_out']
    return st


def test_forced_splitter_repair():
    n, P = 1 << 20, 16
    note("fault: splitters n=%d P=%d" % (n, P))
    st = learned_vs_plain(_lib.FAULT_SPLITTERS, 0, n, P)
    assert st['splitters_exact'] == 0 and st['splitters_repaired'] == P - 1
    assert st['repair_rounds'] >= 2


@pytest.mark.parametrize("mode", ["merged", "exchange"])
def test_forced_bucket_overflow_takes_the_bucket_fallback(mode):
    """a capacity of 300 records on 4 ranks (FirstSort buckets of 256 on
    average): some hundreds of buckets overflow, fewer than the bucket
    fallback's limit"""
    n, P = 1 << 20, 4
    note("fault: bucket cap 300 mode=%s" % mode)
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    st = learned_vs_plain(_lib.FAULT_BUCKET_CAP, 300, n, P, flags=flags)
    assert 0 < st['buckets_overflow'] <= 4096
    assert st['plain_sides'] == 0


def test_forced_bucket_overflow_beyond_the_fallback_limit():
    """a capacity of 8 records: most buckets overflow, more than the bucket
    fallback takes, so the whole side is radix-sorted"""
    n, P = 1 << 20, 16
    note("fault: bucket cap 8")
    st = learned_vs_plain(_lib.FAULT_BUCKET_CAP, 8, n, P)
    assert st['buckets_overflow'] > 4096
    assert st['plain_sides'] >= 1


@pytest.mark.parametrize("mode", ["merged", "exchange"])
def test_forced_plain_fallback(mode):
    n, P = 1 << 20, 16
    note("fault: plain fallback mode=%s" % mode)
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    st = learned_vs_plain(_lib.FAULT_PLAIN, 0, n, P, dist="zipf", flags=flags)
    assert st['plain_sides'] >= 2  # both sides


@pytest.mark.parametrize("P", [1, 8])
@pytest.mark.parametrize("dtype", ["rec16", "u8"])
def test_forced_corruption_is_detected(P, dtype):
    """plain path: a byte past the key flipped between the two local sorts
    (16-byte records, several ranks), or the recorded input sum changed"""
    n = 100000
    g = numpy.random.default_rng(3)
    if dtype == "rec16":
        a = workloads.records16(g.integers(0, 2 ** 64, n, dtype=numpy.uint64))
        field = 'key'
    else:
        a = g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
        field = None
    note("fault: corrupt P=%d dtype=%s" % (P, dtype))
    _lib.check(lib.mpsort_gpu_set_faults(_lib.FAULT_CORRUPT, 0))
    with pytest.raises(_lib.MPSortError) as e:
        gpu_sort(a, even(n, P), even(n, P), field)
    assert e.value.code == _lib.MPSORT_ERR_CHECKSUM
    # without the checksum nothing is checked
    gpu_sort(a, even(n, P), even(n, P), field, flags=_lib.NO_CHECKSUM)


def test_forced_corruption_learned_and_python_face():
    n, P = 1 << 18, 8
    model = trained("uniform", n)
    a = workloads.records16(keys_of("uniform", n, 5))
    _lib.check(lib.mpsort_gpu_set_faults(_lib.FAULT_CORRUPT, 0))
    try:
        with pytest.raises(_lib.MPSortError) as e:
            learned.sort_learned(a, model, 'key', nin=even(n, P))
        assert e.value.code == _lib.MPSORT_ERR_CHECKSUM
        with pytest.raises(_lib.MPSortError):
            mpsort.sort(a.copy(), 'key', comm=mpsort.VirtualComm(vranks=P))
    finally:
        model.close()
