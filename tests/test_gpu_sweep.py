"""Seeded sweep of the GPU paths over sizes, partitions, record types and
key distributions chosen at random (the seed names the case): the plain and
the learned histogram sorts (both SecondSort forms) against the stable sort
of the concatenated input, and the batched segmented sort against each
segment's own stable sort.  The learned path's model comes from another
realisation of the same distribution, so bucket tables do not fit the data
exactly."""
import sys

import numpy
import pytest

from gpu_util import gpu_sort
from mpsort import _lib, learned, segments, workloads

pytestmark = pytest.mark.gpu

DISTS = ["uniform", "clustered", "lognormal", "zipf", "nearly_sorted", "few_values", "one_value", "narrow",
         "two_clumps"]


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def raw_keys(dist, n, g):
    """uint64 keys of a distribution (seeded by the generator g)"""
    s = int(g.integers(1, 1 << 30))
    if dist == "uniform":
        return workloads.uniform(n, s)
    if dist == "clustered":
        return workloads.clustered_keys(n, s)
    if dist == "lognormal":
        return workloads.lognormal_mix(n, s)
    if dist == "zipf":
        return workloads.zipf(n, s)
    if dist == "nearly_sorted":
        return workloads.nearly_sorted(n, s)
    if dist == "few_values":
        return g.integers(0, 5, n).astype(numpy.uint64) * numpy.uint64(0x0123456789ABCDEF)
    if dist == "one_value":
        return numpy.full(n, 0xFEDCBA9876543210, dtype=numpy.uint64)
    if dist == "narrow":
        return numpy.uint64(1 << 40) + g.integers(0, 3000, n).astype(numpy.uint64)
    if dist == "two_clumps":
        c = numpy.where(g.random(n) < 0.5, numpy.uint64(1 << 20), numpy.uint64(1 << 62))
        return c + g.integers(0, 1 << 16, n).astype(numpy.uint64)
    raise ValueError(dist)


def records(kind, keys):
    n = len(keys)
    if kind == "16B":
        a = numpy.zeros(n, workloads.REC16)
        a['key'] = keys
        a['idx'] = numpy.arange(n, dtype=numpy.uint64)
        return a, 'key'
    if kind == "8B":
        a = numpy.zeros(n, workloads.REC8)
        a['key'] = (keys >> numpy.uint64(32)).astype(numpy.uint32).view(numpy.int32)
        a['value'] = numpy.arange(n, dtype=numpy.uint32).view(numpy.float32)
        return a, 'key'
    a = numpy.zeros(n, workloads.PARTICLE32)
    a['id'] = keys
    a['pos'][:, 0] = numpy.arange(n, dtype=numpy.float32)
    return a, 'id'


def partition(n, P, g):
    """P uneven counts summing to n, some of them zero"""
    if P == 1:
        return [n]
    cuts = numpy.sort(g.integers(0, n + 1, P - 1))
    c = numpy.diff(numpy.concatenate([[0], cuts, [n]]))
    return [int(x) for x in c]


def stable_ref(a, field):
    return a[numpy.argsort(a[field], kind='stable')]


@pytest.mark.parametrize("seed", range(36))
def test_histsort_sweep(seed):
    g = numpy.random.default_rng(1000 + seed)
    dist = DISTS[seed % len(DISTS)]
    kind = ["16B", "8B", "32B"][int(g.integers(0, 3))]
    n = int(g.integers(1, 300000)) if seed % 4 else int(g.integers(1, 3000))
    P = int(g.integers(1, 70))
    nin, nout = partition(n, P, g), partition(n, P, g)
    mode = "exchange" if g.random() < 0.3 else "merged"
    note("sweep seed=%d dist=%s kind=%s n=%d P=%d mode=%s" % (seed, dist, kind, n, P, mode))
    a, field = records(kind, raw_keys(dist, n, g))
    ref = stable_ref(a, field)
    plain, _ = gpu_sort(a, nin, nout, field)
    assert numpy.array_equal(plain.view(numpy.uint8), ref.view(numpy.uint8)), "plain path"
    t, tf = records(kind, raw_keys(dist, max(n, 64), g))
    ts, _ = gpu_sort(t, [len(t)], [len(t)], tf)
    model = learned.train(ts, tf, nsample=min(len(t), 1 << 16), seed=seed)
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    out = learned.sort_learned(a, model, field, nin=nin, nout=nout, flags=flags)
    model.close()
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8)), "learned path"


@pytest.mark.parametrize("seed", range(12))
def test_segments_sweep(seed):
    g = numpy.random.default_rng(2000 + seed)
    dist = DISTS[(seed * 5) % len(DISTS)]
    kind = ["16B", "8B"][seed % 2]
    nseg = int(g.integers(1, 400))
    sizes = numpy.exp(g.uniform(0, numpy.log(40000), nseg)).astype(numpy.int64)
    sizes[g.random(nseg) < 0.1] = 0
    n = int(sizes.sum())
    note("segments sweep seed=%d dist=%s kind=%s nseg=%d n=%d" % (seed, dist, kind, nseg, n))
    if kind == "16B":
        a = numpy.zeros(n, workloads.REC16I)
        a['key'] = raw_keys(dist, n, g).view(numpy.int64)
        a['value'] = numpy.arange(n, dtype=numpy.int64)
    else:
        a = numpy.zeros(n, workloads.REC8)
        a['key'] = (raw_keys(dist, n, g) >> numpy.uint64(32)).astype(numpy.uint32).view(numpy.int32)
        a['value'] = numpy.arange(n, dtype=numpy.uint32).view(numpy.float32)
    out = segments.sort_segments(a, sizes, 'key')
    o = 0
    for s in sizes:
        s = int(s)
        seg = a[o:o + s]
        exp = seg[numpy.argsort(seg['key'], kind='stable')]
        assert numpy.array_equal(out[o:o + s].view(numpy.uint8), exp.view(numpy.uint8)), (o, s)
        o += s
