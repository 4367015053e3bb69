"""GPU tests of the learned-splitter path: bit-identical output and counts
against the plain GPU path and the oracle on every key distribution of the
learned workload, forced bucket overflow, the training kernels' gradient
against the CPU oracle, and what the trained network gains."""
import sys

import numpy
import pytest

import ctypes

import oracle
from gpu_util import even, gpu_sort, last_counts
from mpsort import _lib, learned, workloads
from mpsort._lib import lib

pytestmark = pytest.mark.gpu


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def keys_of(dist, n, seed):
    if dist == "uniform":
        return workloads.uniform(n, seed)
    if dist == "clustered":
        return workloads.clustered_keys(n, seed)
    if dist == "lognormal":
        return workloads.lognormal_mix(n, seed)
    if dist == "zipf":
        return workloads.zipf(n, seed)
    if dist == "nearly_sorted":
        return workloads.nearly_sorted(n, seed)
    raise ValueError(dist)


def trained(dist, n, seed=101, linear_root=False, nsample=1 << 16):
    """model from one plain GPU sort of another realisation of dist"""
    t = workloads.records16(keys_of(dist, n, seed))
    ts, _ = gpu_sort(t, [n], [n], 'key')
    return learned.train(ts, 'key', nsample=nsample, seed=7, linear_root=linear_root)


@pytest.mark.parametrize("mode", ["merged", "exchange"])
@pytest.mark.parametrize("dist", ["uniform", "clustered", "lognormal", "zipf", "nearly_sorted"])
def test_learned_equals_plain_and_counts(dist, mode):
    """Both SecondSort forms: buckets gathered from the senders' sorted runs
    (default) and the explicit exchange followed by the bucket sort."""
    n, P = 1 << 20, 16
    note("learned dist=%s n=%d P=%d mode=%s" % (dist, n, P, mode))
    model = trained(dist, n)
    a = workloads.records16(keys_of(dist, n, 5))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    ref_counts = last_counts(P, 8)
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    out = learned.sort_learned(a, model, 'key', nin=even(n, P), flags=flags)
    assert numpy.array_equal(out, ref), "learned output differs from the plain path"
    c = last_counts(P, 8)
    for k in ("C", "P", "myCLT", "myCLE", "myC", "send"):
        assert numpy.array_equal(c[k], ref_counts[k]), k
    st = learned.last_stats() if hasattr(learned, "last_stats") else None
    from mpsort import last_stats
    s = last_stats()
    assert s['learned'] == 1
    note("  exact=%d repaired=%d rounds=%d overflow=%d lds_radix=%d plain_sides=%d" % (
        s['splitters_exact'], s['splitters_repaired'], s['repair_rounds'], s['buckets_overflow'],
        s['buckets_lds_radix'], s['plain_sides']))
    model.close()


@pytest.mark.parametrize("mode", ["merged", "exchange"])
@pytest.mark.parametrize("nvalues", [40, 400])
@pytest.mark.parametrize("shape", ["one", "pairs"])
def test_few_value_buckets_of_a_uniform_model(shape, nvalues, mode):
    """A model of uniform keys and data of few values spread over the key
    range: a bucket then holds one value hundreds or thousands of times
    without being a single-value bucket of the tables (40 values: FirstSort
    buckets; 400: SecondSort buckets).  "one": each value alone, so every
    crowded bin of the in-LDS sort holds one key value and is placed in
    record order there, with no LDS radix sort.  "pairs": next to each value
    v, rare keys v + 1 and v + 2^40 (0.3% each), so v + 1 usually escapes the
    bucket's 64-key sample and shares v's bin: crowded bins of two values,
    which go to the LDS radix sort."""
    n, P = 1 << 20, 16
    note("few-value buckets shape=%s nvalues=%d mode=%s" % (shape, nvalues, mode))
    model = trained("uniform", n)
    g = numpy.random.default_rng(nvalues)
    values = (numpy.arange(nvalues, dtype=numpy.uint64) * numpy.uint64((1 << 64) // nvalues - 12345)
              + numpy.uint64(777))
    keys = values[g.integers(0, nvalues, n)]
    if shape == "pairs":
        u = g.random(n)
        keys = keys + numpy.where(u < 0.003, 1, numpy.where(u < 0.006, 1 << 40, 0)).astype(numpy.uint64)
    a = workloads.records16(keys)
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    out = learned.sort_learned(a, model, 'key', nin=even(n, P), flags=flags)
    from mpsort import last_stats
    s = last_stats()
    note("  overflow=%d lds_radix=%d" % (s['buckets_overflow'], s['buckets_lds_radix']))
    model.close()
    assert numpy.array_equal(out, ref), "learned output differs from the plain path"
    if shape == "one":
        assert s['buckets_lds_radix'] == 0
    else:
        assert s['buckets_lds_radix'] > 0


@pytest.mark.parametrize("P", [1, 3, 8, 64, 100])
def test_learned_against_oracle(P):
    """output bytes and every per-rank count equal the oracle's, with uneven
    output ranks (P = 100: some empty, so C[i] = 0 for the first splitters)"""
    n = 1 << 15
    model = trained("clustered", n, nsample=4096)
    a = workloads.records16(workloads.clustered_keys(n, 9))
    g = numpy.random.default_rng(P)
    cuts = numpy.sort(g.integers(0, n + 1, P - 1))
    nout = list(numpy.diff(numpy.concatenate([[0], cuts, [n]])))
    if P == 100:
        nout = [0, 0, 0] + nout[3:-3] + [nout[0] + nout[1] + nout[2] + sum(nout[-3:]), 0, 0]
    key = oracle.key_for(a.dtype, 'key')
    ref, info = oracle.histogram_sort(a, even(n, P), nout, key)
    out = learned.sort_learned(a, model, 'key', nin=even(n, P), nout=nout)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))
    if P > 1:
        c = last_counts(P, 8)
        for k in ("C", "P", "myCLT", "myCLE", "myC", "send"):
            assert numpy.array_equal(c[k], info[k]), k


@pytest.mark.parametrize("mode", ["merged", "exchange"])
def test_mismatched_model_overflows_and_stays_exact(mode):
    """A model of uniform keys applied to Zipf keys: buckets overflow and fall
    back to the radix path; the output is still the stable sort."""
    n, P = 1 << 20, 4
    model = trained("uniform", n)
    a = workloads.records16(workloads.zipf(n, 3))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    flags = _lib.LEARNED_EXCHANGE if mode == "exchange" else 0
    out = learned.sort_learned(a, model, 'key', nin=even(n, P), flags=flags)
    from mpsort import last_stats
    s = last_stats()
    note("  overflow buckets=%d plain sides=%d" % (s['buckets_overflow'], s['plain_sides']))
    assert s['buckets_overflow'] > 0
    assert numpy.array_equal(out, ref)


@pytest.mark.parametrize("P", [1, 4])
def test_heavy_keys_general_packing(P):
    """Zipf keys through the general (packed key + index) path: single-value
    buckets larger than the in-LDS cap, and at P=1 a sort whose scatter
    source is also its destination."""
    n = 1 << 19
    model = trained("zipf", n)
    a = workloads.records16(workloads.zipf(n, 8))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    out = learned.sort_learned(a, model, 'key', nin=even(n, P), flags=_lib.FORCE_GENERAL)
    assert numpy.array_equal(out, ref)


def test_heavy_keys_in_place_device():
    """Zipf keys sorted in place in device memory (input buffer = output
    buffer) at P=1 and P=4."""
    n = 1 << 19
    model = trained("zipf", n)
    a = workloads.records16(workloads.zipf(n, 12))
    for P in (1, 4):
        ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
        d = lib.mpsort_gpu_dmalloc(a.nbytes)
        try:
            _lib.check(lib.mpsort_gpu_memcpy_h2d(d, a.ctypes.data, a.nbytes))
            nin = (ctypes.c_int64 * P)(*even(n, P))
            key = learned._key_desc(a.dtype, 'key')
            _lib.check(lib.mpsort_gpu_newarray_learned(d, nin, P, d, nin, a.dtype.itemsize, ctypes.byref(key),
                                                       model._h, _lib.DEVICE_PTRS))
            out = numpy.empty_like(a)
            _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, d, a.nbytes))
        finally:
            lib.mpsort_gpu_dfree(d)
        assert numpy.array_equal(out, ref), P


@pytest.mark.parametrize("dtype", ["8B", "32B", "i4"])
def test_learned_record_shapes(dtype):
    n, P = 200000, 5
    g = numpy.random.default_rng(3)
    if dtype == "8B":
        a = numpy.zeros(n, workloads.REC8)
        a['key'] = g.integers(-2 ** 31, 2 ** 31, n)
        a['value'] = numpy.arange(n, dtype='f4')
        f = 'key'
    elif dtype == "32B":
        a = numpy.zeros(n, workloads.PARTICLE32)
        a['id'] = workloads.permutation(n, 4)
        a['pos'] = g.random((n, 3), dtype=numpy.float32)
        f = 'id'
    else:
        a = g.integers(-2 ** 31, 2 ** 31, n).astype('i4')
        f = None
    ts, _ = gpu_sort(a, [n], [n], f)
    model = learned.train(ts, f, nsample=1 << 14)
    ref, _ = gpu_sort(a, even(n, P), even(n, P), f)
    out = learned.sort_learned(a, model, f, nin=even(n, P))
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))


def test_training_gradient_matches_cpu_oracle():
    """Gradient of the root's loss from the GPU training kernels (FP32)
    against the double-precision oracle restatement."""
    g = numpy.random.default_rng(11)
    raw = g.normal(0, 0.7, oracle.MLP_PARAMS).astype('f4')
    x = numpy.sort(g.random(5000)).astype('f4')
    y = numpy.sort(g.random(5000)).astype('f4')
    loss_g, grad_g = learned.mlp_grad(raw, x, y)
    loss_c, grad_c = oracle.mlp_grad(raw.astype('f8'), x, y)
    assert abs(loss_g - loss_c) <= 1e-4 * abs(loss_c) + 1e-6
    scale = numpy.abs(grad_c).max()
    assert numpy.all(numpy.abs(grad_g - grad_c) <= 2e-4 * scale + 1e-4 * numpy.abs(grad_c)), \
        numpy.abs(grad_g - grad_c).max()


def power_keys(n, seed, power=4):
    """keys 2^63 * u^power, u uniform on [0, 1): a smooth CDF that is far from
    linear in key space and in log space alike (and almost no duplicates)"""
    u = (workloads.uniform(n, seed) >> numpy.uint64(11)).astype(numpy.float64) * 2.0 ** -53
    return (u ** power * 2.0 ** 63).astype(numpy.uint64)


@pytest.mark.parametrize("dist", ["power", "lognormal"])
def test_trained_network_gain(dist):
    """What the trained network buys over the same two-level index with a
    linear root (the method without the network), on keys whose CDF is
    smooth but non-linear in the root's input: a much smaller root error,
    and on the power-law keys (no duplicates) far fewer samples in the
    fullest leaf, a smaller model error and no more overflowing buckets."""
    n, P = 1 << 20, 16
    keys = (lambda seed: power_keys(n, seed)) if dist == "power" else (lambda seed: workloads.lognormal_mix(n, seed))

    def fit(linear_root):
        t = workloads.records16(keys(101))
        ts, _ = gpu_sort(t, [n], [n], 'key')
        return learned.train(ts, 'key', nsample=1 << 16, seed=7, linear_root=linear_root)

    m_net, m_lin = fit(False), fit(True)
    i_net, i_lin = m_net.info(), m_lin.info()
    note("  %s net: max_leaf=%d err=%.2e loss=%.2e steps=%d log=%d | linear: max_leaf=%d err=%.2e loss=%.2e log=%d" % (
        dist, i_net['max_leaf_samples'], i_net['err_max'], i_net['loss'], i_net['steps'], i_net['log_input'],
        i_lin['max_leaf_samples'], i_lin['err_max'], i_lin['loss'], i_lin['log_input']))
    assert i_net['steps'] > 0 and i_lin['steps'] == 0
    assert i_net['loss'] * 10 < i_lin['loss']
    a = workloads.records16(keys(5))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    from mpsort import last_stats
    out = learned.sort_learned(a, m_net, 'key', nin=even(n, P))
    s_net = last_stats()
    assert numpy.array_equal(out, ref)
    out = learned.sort_learned(a, m_lin, 'key', nin=even(n, P))
    s_lin = last_stats()
    assert numpy.array_equal(out, ref)
    note("  overflow buckets: net %d linear %d; lds radix buckets: net %d linear %d" % (
        s_net['buckets_overflow'], s_lin['buckets_overflow'], s_net['buckets_lds_radix'],
        s_lin['buckets_lds_radix']))
    if dist == "power":
        assert i_net['max_leaf_samples'] * 2 < i_lin['max_leaf_samples']
        assert i_net['err_max'] < i_lin['err_max']
        assert s_net['buckets_overflow'] <= s_lin['buckets_overflow']


def test_model_is_monotone():
    n = 1 << 18
    m = trained("clustered", n, nsample=1 << 14)
    k = numpy.sort(workloads.uniform(200000, 3) >> numpy.uint64(1))
    f = m.eval(k)
    assert numpy.all(numpy.diff(f) >= 0)


def test_merged_secondsort_with_thousands_of_oversized_buckets():
    """2^26 records on 2 virtual ranks: every receiver bucket of the merged
    SecondSort holds about 10,900 records, above the largest in-LDS bucket,
    so more buckets overflow than the bucket fallback lists; the receivers
    are then radix-sorted whole.  Output equals the plain path's."""
    n, P = 1 << 26, 2
    note("merged overflow n=%d P=%d" % (n, P))
    model = trained("uniform", 1 << 22)
    a = workloads.records16(keys_of("uniform", n, 9))
    ref, _ = gpu_sort(a, even(n, P), even(n, P), 'key')
    out = learned.sort_learned(a, model, 'key', nin=even(n, P))
    st = _lib.last_stats()
    note("overflow buckets %d, plain sides %d" % (st.buckets_overflow, st.plain_sides))
    assert st.buckets_overflow > 4096
    assert st.plain_sides >= 1
    assert numpy.array_equal(out, ref)
    assert st.checksum_in == st.checksum_out
