"""Parity at the workloads' stated sizes, through properties that do not need
a CPU sort of the whole array:

- config 2: 2^28 x 16 B clustered Morton records on 64 virtual ranks, plain
  path: sorted, stable (ties in input order), a permutation, checksums equal,
  and every per-rank count against its closed form computed on the host from
  the sorted output (P_i = k[C[i+1]-1]; #{x < P_i} and #{x <= P_i} of each
  sender; the send matrix from the stable order), which the oracle tests pin
  on small inputs (tests/test_oracle.py);
- config 3 (a)-(d): the learned path at 2^28 on each distribution equals the
  plain path byte for byte, with the same counts;
- config 4: 10,000 independent segments, sizes log-uniform on [1e3, 1e7]
  rescaled to 2^30 records in total, {i32, f32} and {i64, i64} records, in one
  graph-captured tuned plan: every segment sorted, stable and a permutation of
  itself."""
import ctypes
import sys

import numpy
import pytest

from gpu_util import even, last_counts
from mpsort import _lib, learned, segments, workloads
from mpsort._lib import lib

pytestmark = pytest.mark.gpu

N28 = 1 << 28
P64 = 64


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sort_device(recs, P, model=None):
    """plain (or learned) sort of host records through device buffers;
    returns (sorted host copy, stats, counts)"""
    n = len(recs)
    din = lib.mpsort_gpu_dmalloc(recs.nbytes)
    dout = lib.mpsort_gpu_dmalloc(recs.nbytes)
    assert din and dout
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(din, recs.ctypes.data, recs.nbytes))
        c = (ctypes.c_int64 * P)(*even(n, P))
        key = _lib.KeyDesc(offset=0, width=8, is_signed=0, nwords=1, reserved=0)
        if model is None:
            rc = lib.mpsort_gpu_newarray(din, c, P, dout, c, 16, ctypes.byref(key), _lib.DEVICE_PTRS)
        else:
            rc = lib.mpsort_gpu_newarray_learned(din, c, P, dout, c, 16, ctypes.byref(key), model._h,
                                                 _lib.DEVICE_PTRS)
        _lib.check(rc)
        st = _lib.last_stats()
        cnt = last_counts(P, 8)
        out = numpy.empty_like(recs)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, dout, recs.nbytes))
    finally:
        lib.mpsort_gpu_dfree(din)
        lib.mpsort_gpu_dfree(dout)
    return out, st, cnt


def _check_stable_permutation(out, n):
    k = out['key']
    assert bool(numpy.all(k[1:] >= k[:-1])), "not sorted"
    ties = k[1:] == k[:-1]
    idx = out['idx'].astype(numpy.int64)
    assert bool(numpy.all(idx[1:][ties] > idx[:-1][ties])), "equal keys out of input order"
    del ties
    seen = numpy.zeros(n, dtype=bool)
    seen[idx] = True
    assert bool(seen.all()), "not a permutation"


def _counts_before(sender, positions, P):
    """for each position b (non-decreasing), the number of records of each
    sender among output positions [0, b): a (len(positions), P) array"""
    res = numpy.zeros((len(positions), P), numpy.int64)
    acc = numpy.zeros(P, numpy.int64)
    prev = 0
    for i, b in enumerate(positions):
        if b > prev:
            acc += numpy.bincount(sender[prev:b], minlength=P)
            prev = b
        res[i] = acc
    return res


def _closed_form_counts(out, n, P):
    """C, splitters, myCLT, myCLE, myC and send of the reference's histogram
    sort (mpsort-mpi.c:333-603), from the globally sorted output (even input
    and output partitions)"""
    k = out['key']
    per = n // P
    sender = (out['idx'].astype(numpy.int64) // per).astype(numpy.int16)
    C = numpy.array([n * t // P for t in range(P + 1)], numpy.int64)
    split = k[C[1:P] - 1]
    lo = numpy.searchsorted(k, split, 'left')
    hi = numpy.searchsorted(k, split, 'right')
    clt = numpy.zeros((P, P + 1), numpy.int64)
    cle = numpy.zeros((P, P + 1), numpy.int64)
    myc = numpy.zeros((P, P + 1), numpy.int64)
    clt[:, 1:P] = _counts_before(sender, lo, P).T
    cle[:, 1:P] = _counts_before(sender, hi, P).T
    myc[:, 1:P] = _counts_before(sender, C[1:P], P).T
    nj = numpy.bincount(sender, minlength=P)
    clt[:, P] = cle[:, P] = myc[:, P] = nj
    send = numpy.diff(myc, axis=1)
    return dict(C=C, split=split, myCLT=clt, myCLE=cle, myC=myc, send=send)


def test_config2_plain_full_size_counts_closed_form():
    n, P = N28, P64
    recs = workloads.clustered_records(n, seed=2)
    note("config 2 plain n=%d P=%d" % (n, P))
    out, st, c = _sort_device(recs, P)
    assert st.checksum_in == st.checksum_out
    del recs
    _check_stable_permutation(out, n)
    want = _closed_form_counts(out, n, P)
    assert numpy.array_equal(c['C'], want['C'])
    assert numpy.array_equal(c['P'].view(numpy.uint64).reshape(-1), want['split']), "splitters"
    for name in ("myCLT", "myCLE", "myC", "send"):
        assert numpy.array_equal(c[name], want[name]), name
    note("config 2: %d bisection rounds, phases first %.2f findP %.2f exch %.2f second %.2f ms"
         % (st.bisect_rounds, st.t_first_sort_ms, st.t_findp_ms, st.t_exchange_ms, st.t_second_sort_ms))


GENS = {"uniform": workloads.uniform, "lognormal": workloads.lognormal_mix, "zipf": workloads.zipf,
        "nearly_sorted": workloads.nearly_sorted}


@pytest.mark.parametrize("dist", ["uniform", "lognormal", "zipf", "nearly_sorted"])
def test_config3_learned_full_size_equals_plain(dist):
    n, P = N28, P64
    g = GENS[dist]
    note("config 3 %s: training realisation" % dist)
    t = workloads.records16(g(1 << 24, 101))
    ts, _, _ = _sort_device(t, 1)
    model = learned.train(ts, 'key', nsample=1 << 20, seed=1)
    del t, ts
    recs = workloads.records16(g(n, 5))
    note("config 3 %s: plain n=%d P=%d" % (dist, n, P))
    ref, st_p, c_p = _sort_device(recs, P)
    note("config 3 %s: learned" % dist)
    got, st_l, c_l = _sort_device(recs, P, model)
    model.close()
    del recs
    assert numpy.array_equal(got, ref), "learned output differs from the plain path"
    del got
    _check_stable_permutation(ref, n)
    for name in ("C", "P", "myCLT", "myCLE", "myC", "send"):
        assert numpy.array_equal(c_l[name], c_p[name]), name
    assert st_l.checksum_in == st_l.checksum_out == st_p.checksum_out
    note("config 3 %s: splitters exact %d repaired %d, repair rounds %d (plain bisection %d), overflow buckets %d, "
         "plain sides %d, LDS radix buckets %d" % (dist, st_l.splitters_exact, st_l.splitters_repaired,
                                                   st_l.repair_rounds, st_p.bisect_rounds, st_l.buckets_overflow,
                                                   st_l.plain_sides, st_l.buckets_lds_radix))


def test_config4_batched_full_size():
    nseg, total = 10000, 1 << 30
    sizes = workloads.segment_sizes(nseg, 4, lo=1e3, hi=1e7, total=total).astype(numpy.int64)
    # rescaled to 2^30 in total by flooring (each at least 1 record)
    assert abs(int(sizes.sum()) - total) < nseg and sizes.min() >= 1
    total = int(sizes.sum())
    pick = numpy.array([workloads.mix64(0xB0A7 + i) & 1 for i in range(nseg)], dtype=bool)
    plan = segments.SegmentPlan(graph=True, tune=True)
    groups = []
    try:
        for gi, (dt, sel) in enumerate(((workloads.REC8, ~pick), (workloads.REC16I, pick))):
            sz = sizes[sel]
            n = int(sz.sum())
            a = numpy.zeros(n, dt)
            raw = workloads.uniform(n, 40 + gi)
            if dt == workloads.REC8:
                a['key'] = (raw >> numpy.uint64(32)).astype(numpy.uint32).view(numpy.int32)
                a['value'] = numpy.arange(n, dtype=numpy.uint32).view(numpy.float32)
            else:
                a['key'] = raw.view(numpy.int64)
                a['value'] = numpy.arange(n, dtype=numpy.int64)
            del raw
            d = lib.mpsort_gpu_dmalloc(a.nbytes)
            assert d
            _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, a.nbytes))
            groups.append((dt, sz, n, d, a.nbytes))
            plan.add(d, d, sz, dt, 'key')
            del a
        note("config 4: %d segments, %d records, groups %s" % (nseg, total, [g[2] for g in groups]))
        plan.run()
        note("config 4: plan %s" % [plan.info(i) for i in range(2)])
        for dt, sz, n, d, nb in groups:
            out = numpy.empty(n, dt)
            _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, d, nb))
            starts = numpy.concatenate([[0], numpy.cumsum(sz)[:-1]])
            seg_of = numpy.repeat(numpy.arange(len(sz), dtype=numpy.int32), sz)
            k = out['key']
            v = (out['value'].view(numpy.uint32) if dt == workloads.REC8 else out['value']).astype(numpy.int64)
            same = seg_of[1:] == seg_of[:-1]
            assert bool(numpy.all(~same | (k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (v[1:] > v[:-1])))), \
                "a segment is not stably sorted"
            del same, k
            assert bool(numpy.all((v >= starts[seg_of]) & (v < (starts + sz)[seg_of]))), "a record left its segment"
            seen = numpy.zeros(n, dtype=bool)
            seen[v] = True
            assert bool(seen.all()), "records lost"
            del out, seg_of, seen, v
    finally:
        plan.close()
        for g in groups:
            lib.mpsort_gpu_dfree(g[3])
