"""GPU tests of the batched segmented sort (many independent stable sorts):
every segment equals its own stable sort, for the two record types of the
batched workload ({i32 key, f32 value}, {i64 key, i64 value}), small
segments (in LDS) and large ones (onesweep), duplicate-heavy keys, empty
segments, graph replays against eager runs, and the autotuned plan."""
import ctypes

import numpy
import pytest

from mpsort import _lib, segments, workloads
from mpsort._lib import lib

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)


def make(dtype, sizes, seed, key_bits=None):
    n = int(sum(sizes))
    a = numpy.zeros(n, dtype)
    raw = workloads.uniform(n, seed)
    if key_bits is not None:
        raw = raw >> numpy.uint64(64 - key_bits)
    if a['key'].dtype == numpy.int32:
        a['key'] = (raw >> numpy.uint64(32)).astype(numpy.uint32).view(numpy.int32)
        a['value'] = numpy.arange(n, dtype=numpy.float32)
    else:
        a['key'] = raw.view(numpy.int64)
        a['value'] = numpy.arange(n, dtype=numpy.int64)
    return a


def expected(a, sizes):
    out = numpy.empty_like(a)
    o = 0
    for s in sizes:
        s = int(s)
        seg = a[o:o + s]
        out[o:o + s] = seg[numpy.argsort(seg['key'], kind='stable')]
        o += s
    return out


def sizes_mixed(nseg, seed, total):
    s = workloads.segment_sizes(nseg, seed, lo=1e1, hi=1e5, total=total).astype(numpy.int64)
    s[::17] = 0  # some empty segments
    return s


@pytest.mark.parametrize("dtype", [workloads.REC8, workloads.REC16I], ids=["8B", "16B"])
def test_segments_equal_per_segment_stable_sort(dtype):
    sizes = sizes_mixed(300, 3, 1 << 21)
    a = make(dtype, sizes, 11)
    note("segments: %d records %d" % (len(sizes), int(sizes.sum())))
    out = segments.sort_segments(a, sizes, 'key')
    assert numpy.array_equal(out, expected(a, sizes))


@pytest.mark.parametrize("dtype", [workloads.REC8, workloads.REC16I], ids=["8B", "16B"])
def test_segments_duplicate_heavy(dtype):
    """keys drawn from 2^6 values: crowded bins, single-value segments"""
    sizes = sizes_mixed(120, 5, 1 << 20)
    a = make(dtype, sizes, 12, key_bits=6)
    out = segments.sort_segments(a, sizes, 'key')
    assert numpy.array_equal(out, expected(a, sizes))


def test_segments_graph_replays_match_eager():
    """A plan of both record types captured once and replayed three times,
    with eager work and synchronisation in between; every replay equals the
    eager result."""
    s8 = sizes_mixed(200, 7, 1 << 20)
    s16 = sizes_mixed(150, 8, 1 << 20)
    a8 = make(workloads.REC8, s8, 21)
    a16 = make(workloads.REC16I, s16, 22)
    e8, e16 = expected(a8, s8), expected(a16, s16)
    d8 = lib.mpsort_gpu_dmalloc(a8.nbytes)
    d16 = lib.mpsort_gpu_dmalloc(a16.nbytes)
    o16 = lib.mpsort_gpu_dmalloc(a16.nbytes)
    try:
        plan = segments.SegmentPlan(graph=True)
        plan.add(d8, d8, s8, workloads.REC8, 'key')      # in place
        plan.add(d16, o16, s16, workloads.REC16I, 'key')  # out of place
        for rep in range(3):
            _lib.check(lib.mpsort_gpu_memcpy_h2d(d8, a8.ctypes.data, a8.nbytes))
            _lib.check(lib.mpsort_gpu_memcpy_h2d(d16, a16.ctypes.data, a16.nbytes))
            # eager work of the program between replays
            eager = segments.sort_segments(a16[:100000], [100000], 'key')
            assert numpy.array_equal(eager, expected(a16[:100000], [100000]))
            ms = plan.run()
            _lib.check(lib.mpsort_gpu_synchronize())
            g8 = numpy.empty_like(a8)
            g16 = numpy.empty_like(a16)
            _lib.check(lib.mpsort_gpu_memcpy_d2h(g8.ctypes.data, d8, a8.nbytes))
            _lib.check(lib.mpsort_gpu_memcpy_d2h(g16.ctypes.data, o16, a16.nbytes))
            note("replay %d: %.3f ms" % (rep, ms))
            assert numpy.array_equal(g8, e8), rep
            assert numpy.array_equal(g16, e16), rep
        plan.close()
    finally:
        for p in (d8, d16, o16):
            lib.mpsort_gpu_dfree(p)


def test_segments_tuned_plan(tmp_path, monkeypatch):
    monkeypatch.setenv("MPSORT_GPU_TUNE_CACHE", str(tmp_path / "tune.txt"))
    sizes = sizes_mixed(100, 9, 1 << 20)
    a = make(workloads.REC16I, sizes, 31)
    d = lib.mpsort_gpu_dmalloc(a.nbytes)
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, a.nbytes))
        plan = segments.SegmentPlan(tune=True)
        plan.add(d, d, sizes, a.dtype, 'key')
        # tuning runs candidates in place on the same buffer: start again
        plan.run()
        _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, a.nbytes))
        plan.run()
        info = plan.info()
        note("tuned: %s" % info)
        assert info["block"] in (256, 512, 1024) and 0 <= info["threshold"] <= 9216
        g = numpy.empty_like(a)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(g.ctypes.data, d, a.nbytes))
        plan.close()
    finally:
        lib.mpsort_gpu_dfree(d)
    assert numpy.array_equal(g, expected(a, sizes))


def test_32_byte_records_default_and_tuned():
    """32-byte records: the default tile must be one the digit passes take
    for 8-word records (plain run and the tuned plan)"""
    dt = numpy.dtype([('key', 'i8'), ('value', 'i8'), ('pad', 'u8', 2)])
    sizes = numpy.array([0, 5, 3000, 70000, 1, 250000, 12345], dtype=numpy.int64)
    n = int(sizes.sum())
    g = numpy.random.default_rng(71)
    a = numpy.zeros(n, dt)
    a['key'] = g.integers(-2 ** 63, 2 ** 63 - 1, n, dtype=numpy.int64)
    a['value'] = numpy.arange(n)
    note("segments 32B n=%d" % n)
    want = expected(a, sizes)
    for tune in (False, True):
        out = segments.sort_segments(a, sizes, orderby='key', tune=tune)
        assert numpy.array_equal(out, want), tune


def _plan_sort(a, sizes, form, graph=False, each=False):
    d = lib.mpsort_gpu_dmalloc(max(a.nbytes, 1))
    try:
        _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, a.nbytes))
        plan = segments.SegmentPlan(graph=graph, small_form=form, each=each)
        plan.add(d, d, sizes, a.dtype, 'key')
        plan.run()
        info = plan.info()
        g = numpy.empty_like(a)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(g.ctypes.data, d, a.nbytes))
        plan.close()
    finally:
        lib.mpsort_gpu_dfree(d)
    return g, info


@pytest.mark.parametrize("key_bits", [None, 6], ids=["spread", "dup6"])
@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("dtype", [workloads.REC8, workloads.REC16I], ids=["8B", "16B"])
def test_segments_small_forms(dtype, form, key_bits):
    """the in-LDS segments by 512 x 9 (two workgroups per CU) or 512 x 6
    (three per CU) workgroups, those above the form's capacity by the
    1024-thread form; duplicate-heavy keys take the LDS radix sort in the
    512 x 6 form"""
    sizes = sizes_mixed(200, 13 + form, 1 << 20)
    a = make(dtype, sizes, 41 + form, key_bits=key_bits)
    note("segments form %d keybits %s n=%d" % (form, key_bits, int(sizes.sum())))
    g, info = _plan_sort(a, sizes, form)
    assert info["small_form"] == form and info["small_waves_per_cu"] == (24 if form == 2 else 16)
    assert numpy.array_equal(g, expected(a, sizes))


def test_segments_small_form_graph_replays():
    sizes = sizes_mixed(150, 23, 1 << 20)
    a = make(workloads.REC8, sizes, 51)
    g, _ = _plan_sort(a, sizes, 2, graph=True)
    assert numpy.array_equal(g, expected(a, sizes))


@pytest.mark.parametrize("form", [0, 2])
@pytest.mark.parametrize("dtype", [workloads.REC8, workloads.REC16I], ids=["8B", "16B"])
def test_segments_one_at_a_time(dtype, form):
    """MPSORT_GPU_SEGPLAN_EACH: every segment by launches of its own (the
    per-segment loop the batched plan replaces), small ones in LDS and large
    ones by one-segment views of the radix plan, duplicate-heavy keys too"""
    sizes = sizes_mixed(120, 61 + form, 1 << 20)
    for key_bits in (None, 6):
        a = make(dtype, sizes, 71 + form, key_bits=key_bits)
        note("segments one at a time form %d keybits %s n=%d" % (form, key_bits, int(sizes.sum())))
        g, _ = _plan_sort(a, sizes, form, each=True)
        assert numpy.array_equal(g, expected(a, sizes))
