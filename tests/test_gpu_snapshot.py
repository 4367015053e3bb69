"""GPU tests of the snapshot-matching workload: two snapshots of 32-byte
particles ({u64 id, f32 pos[3], f32 vel[3]}), A with the ids a seeded
permutation and B a copy of A with 0.1% of the records displaced, both
sorted by id on 64 virtual ranks with learned splitters (one model, trained
on A's plain sort, serves both).  After the sort the ids of A and B match
elementwise (they are 0..n-1), the learned output equals the plain GPU
output byte for byte, and, at a size the CPU finishes quickly, equals the
reference's own mpsort_omp (compiled from its sources) byte for byte."""
import ctypes
import sys

import numpy
import pytest

import oracle
from gpu_util import even
from mpsort import _lib, learned, workloads
from mpsort._lib import lib

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    """device copies of host arrays, freed at exit"""

    def __init__(self):
        self.ptrs = []

    def put(self, a):
        p = lib.mpsort_gpu_dmalloc(a.nbytes)
        assert p, "device allocation failed"
        self.ptrs.append(p)
        _lib.check(lib.mpsort_gpu_memcpy_h2d(p, a.ctypes.data, a.nbytes))
        return p

    def empty(self, nbytes):
        p = lib.mpsort_gpu_dmalloc(nbytes)
        assert p, "device allocation failed"
        self.ptrs.append(p)
        return p

    def get(self, p, like):
        out = numpy.empty_like(like)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, p, out.nbytes))
        return out

    def close(self):
        for p in self.ptrs:
            lib.mpsort_gpu_dfree(p)
        self.ptrs = []


def sort_pair(n, P, seed):
    """plain sort of A, model from it, learned sorts of A and B; returns the
    host copies (plain A, learned A, learned B) and the learned statistics"""
    a, b = workloads.snapshot_pair(n, seed)
    key = learned._key_desc(a.dtype, 'id')
    nin = (ctypes.c_int64 * P)(*even(n, P))
    d = Dev()
    try:
        da, db = d.put(a), d.put(b)
        dout = d.empty(a.nbytes)
        _lib.check(lib.mpsort_gpu_newarray(da, nin, P, dout, nin, 32, ctypes.byref(key), _lib.DEVICE_PTRS))
        plain = d.get(dout, a)
        note("snapshot n=%d P=%d: plain sort done" % (n, P))
        model = learned.train(None, nsample=min(n, 1 << 20), seed=1, device_ptr=(dout, n, 32, key))
        outs, stats = [], []
        for src in (da, db):
            _lib.check(lib.mpsort_gpu_newarray_learned(src, nin, P, dout, nin, 32, ctypes.byref(key), model._h,
                                                       _lib.DEVICE_PTRS))
            stats.append(_lib.last_stats().as_dict())
            outs.append(d.get(dout, a))
        note("snapshot n=%d P=%d: learned sorts of A and B done" % (n, P))
        model.close()
    finally:
        d.close()
    return a, b, plain, outs[0], outs[1], stats


def test_snapshot_matching_full_size():
    """2 x 2^27 records of 32 B (4 GiB each), 64 virtual ranks."""
    n, P = 1 << 27, 64
    a, b, plain, la, lb, stats = sort_pair(n, P, 27)
    ids = numpy.arange(n, dtype=numpy.uint64)
    assert numpy.array_equal(plain['id'], ids)
    assert numpy.array_equal(la['id'], ids)
    assert numpy.array_equal(lb['id'], ids), "ids of the two snapshots differ after the sort"
    assert numpy.array_equal(la.view(numpy.uint8), plain.view(numpy.uint8))
    # unique ids: both snapshots hold the same records, so the sorted arrays
    # are identical
    assert numpy.array_equal(lb.view(numpy.uint8), plain.view(numpy.uint8))
    for s in stats:
        assert s["checksum_in"] == s["checksum_out"]
        note("learned: %.3f ms, repaired splitters %d, overflow buckets %d" %
             (s["t_total_ms"], s["splitters_repaired"], s["buckets_overflow"]))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_snapshot_matching_against_reference():
    """2^20 records: the learned output of both snapshots equals the
    reference mpsort_omp output byte for byte."""
    n, P = 1 << 20, 64
    a, b, plain, la, lb, _ = sort_pair(n, P, 5)
    key = oracle.key_for(a.dtype, 'id')
    ra = oracle.ref_mpsort_omp(a, key, nthreads=8)
    rb = oracle.ref_mpsort_omp(b, key, nthreads=8)
    assert numpy.array_equal(la.view(numpy.uint8), ra.view(numpy.uint8))
    assert numpy.array_equal(lb.view(numpy.uint8), rb.view(numpy.uint8))
    assert numpy.array_equal(plain.view(numpy.uint8), ra.view(numpy.uint8))
