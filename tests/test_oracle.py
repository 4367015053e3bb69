"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each oracle function is checked against something it does not share code
with: closed forms and brute force (Python integers, numpy's stable sort),
and the reference's own C compiled from the reference tree (oracle/_ref):
radix_sort, mpsort_omp and, when MPI is present, mpsort_mpi_newarray run on
real MPI ranks.  A plausible mistake (dropped term, wrong sign, off-by-one in a
search, unstable merge, swapped operands) fails at least one of these.
"""
import itertools

import numpy
import pytest

import oracle
from oracle import OKey


def rng(seed):
    return numpy.random.default_rng(seed)


# ---- radix mapping (binding.pyx:81-121) -----------------------------------
@pytest.mark.parametrize("dt,vals,expect", [
    ('u8', [0, 1, 2 ** 64 - 1], [0, 1, 2 ** 64 - 1]),
    ('i8', [-2 ** 63, -1, 0, 2 ** 63 - 1], [0, 2 ** 63 - 1, 2 ** 63, 2 ** 64 - 1]),
    ('u4', [0, 1, 2 ** 32 - 1], [0, 1, 2 ** 32 - 1]),
    ('i4', [-2 ** 31, -1, 0, 2 ** 31 - 1],
     [2 ** 63 - 2 ** 31, 2 ** 63 - 1, 2 ** 63, 2 ** 63 + 2 ** 31 - 1]),
])
def test_radix_edge_vectors(dt, vals, expect):
    a = numpy.array(vals, dtype=dt)
    r = oracle.radix(a, oracle.key_for(a.dtype))
    got = [int.from_bytes(bytes(row), 'little') for row in r]
    assert got == expect


def test_radix_multiword_last_word_most_significant():
    dt = numpy.dtype([('v', 'u4'), ('k', ('i8', 2))])
    a = numpy.zeros(3, dt)
    a['k'] = [[5, -1], [7, 0], [-3, 0]]
    r = oracle.radix(a, oracle.key_for(dt, 'k'))
    ints = [int.from_bytes(bytes(row), 'little') for row in r]
    # word 1 is the most significant: (-1, *) < (0, *) and within (0,*): -3 < 7
    assert ints[0] < ints[2] < ints[1]
    assert r.shape == (3, 16)


# ---- compare / bisect (radixsort.c:47-144) --------------------------------
def test_compar_matches_bigint_order():
    g = rng(1)
    for rs in (2, 4, 8, 12, 16, 24):
        for _ in range(200):
            a = g.integers(0, 256, rs, dtype=numpy.uint8)
            b = a.copy() if g.random() < 0.2 else g.integers(0, 256, rs, dtype=numpy.uint8)
            if g.random() < 0.3:
                b[: rs // 2] = a[: rs // 2]
            ia, ib = int.from_bytes(bytes(a), 'little'), int.from_bytes(bytes(b), 'little')
            assert oracle.compar(a, b) == (ia > ib) - (ia < ib)


def test_bisect_is_floor_mean():
    g = rng(2)
    for rs in (2, 4, 8, 16, 24):
        for _ in range(200):
            a = g.integers(0, 256, rs, dtype=numpy.uint8)
            b = g.integers(0, 256, rs, dtype=numpy.uint8)
            ia, ib = int.from_bytes(bytes(a), 'little'), int.from_bytes(bytes(b), 'little')
            if ia > ib:
                a, b, ia, ib = b, a, ib, ia
            m = int.from_bytes(bytes(oracle.bisect(a, b)), 'little')
            assert m == (ia + ib) // 2
    # the carry case: both at the maximum
    a = numpy.full(8, 255, numpy.uint8)
    assert int.from_bytes(bytes(oracle.bisect(a, a)), 'little') == 2 ** 64 - 1


# ---- stable local sort (radixsort.c:35-44, stdlib/msort.c) ----------------
def _recs(n, seed, nkeys=None):
    g = rng(seed)
    dt = numpy.dtype([('key', 'u8'), ('idx', 'u8')])
    a = numpy.zeros(n, dt)
    a['key'] = g.integers(0, nkeys, n) if nkeys else g.integers(0, 2 ** 64, n, dtype=numpy.uint64)
    a['idx'] = numpy.arange(n)
    return a


@pytest.mark.parametrize("nkeys", [None, 7, 1000])
def test_radix_sort_is_numpy_stable_sort(nkeys):
    a = _recs(5000, 3, nkeys)
    k = oracle.key_for(a.dtype, 'key')
    s = oracle.radix_sort(a, k)
    order = numpy.argsort(a['key'], kind='stable')
    assert numpy.array_equal(s, a[order])


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built (make -C oracle ref)")
@pytest.mark.parametrize("dt,field", [('i4', None), ('u8', None), ('i8', None),
                                      ([('v', 'u4'), ('k', ('u8', 2)), ('e', 'u8')], 'k')])
def test_radix_sort_matches_compiled_reference(dt, field):
    g = rng(4)
    dtype = numpy.dtype(dt)
    a = numpy.zeros(3000, dtype)
    raw = g.integers(0, 2 ** 63, (3000, 4), dtype=numpy.int64)
    b = a.view(numpy.uint8).reshape(3000, -1)
    b[...] = raw.view(numpy.uint8)[:, : dtype.itemsize]
    if field:
        a[field][:, 1] %= 5  # ties in the most significant word
    k = oracle.key_for(dtype, field)
    assert numpy.array_equal(oracle.radix_sort(a, k).view(numpy.uint8),
                             oracle.ref_radix_sort(a, k).view(numpy.uint8))


# ---- histogram sort -------------------------------------------------------
def _check_against_closed_form(a, key, nin, nout):
    out, info = oracle.histogram_sort(a, nin, nout, key)
    assert info['rc'] == 0
    r = oracle.radix(a, key)
    cf = oracle.closed_form(r, nin, nout)
    assert numpy.array_equal(out.view(numpy.uint8), a[cf['order']].view(numpy.uint8))
    P = len(nin)
    got = [int.from_bytes(bytes(row), 'little') for row in info['P']]
    assert got == cf['splitters'][:P - 1]
    assert numpy.array_equal(info['C'], cf['C'])
    assert numpy.array_equal(info['myCLT'], cf['myCLT'])
    assert numpy.array_equal(info['myCLE'], cf['myCLE'])
    assert numpy.array_equal(info['send'], cf['send'])
    assert numpy.array_equal(numpy.diff(info['myC'], axis=1), info['send'])
    assert info['checksum_in'] == info['checksum_out']
    return out, info


@pytest.mark.parametrize("P,nkeys", [(1, None), (2, None), (4, None), (4, 7), (7, 1000), (16, 3)])
def test_histogram_sort_closed_form(P, nkeys):
    a = _recs(1200, 10 + P, nkeys)
    key = oracle.key_for(a.dtype, 'key')
    n = len(a)
    nin = [n * (t + 1) // P - n * t // P for t in range(P)]
    g = rng(P)
    cuts = numpy.sort(g.integers(0, n + 1, P - 1))
    nout = list(numpy.diff(numpy.concatenate([[0], cuts, [n]])))
    _check_against_closed_form(a, key, nin, nout)


def test_histogram_sort_empty_ranks_and_zero_outputs():
    a = _recs(1000, 5, 50)
    key = oracle.key_for(a.dtype, 'key')
    # test_mpsort.py:125-140 partitions
    _check_against_closed_form(a, key, [0, 400, 0, 600], [200, 200, 0, 600])
    _check_against_closed_form(a, key, [1000, 0, 0, 0], [0, 0, 0, 1000])
    _check_against_closed_form(a, key, [0, 0, 0, 1000], [1000, 0, 0, 0])


def test_histogram_sort_size_mismatch():
    a = _recs(10, 1)
    _, info = oracle.histogram_sort(a, [5, 5], [5, 4], oracle.key_for(a.dtype, 'key'))
    assert info['rc'] == -1


def test_few_items_enumeration():
    """test_mpsort.py:330-351: every size vector in {0,1,2}^4, 3-word u8 key,
    36-byte records; the output equals numpy's sort of the keys."""
    dt = numpy.dtype([('vkey', ('u8', 3)), ('vector', ('u4', 3))])
    key = oracle.key_for(dt, 'vkey')
    for sizes in itertools.product(*([[0, 1, 2]] * 4)):
        parts = []
        for t in range(4):
            s = numpy.zeros(sizes[t], dt)
            s['vkey'] = numpy.arange(sizes[t], dtype='u8')[:, None]
            s['vector'] = 1
            parts.append(s)
        S = numpy.concatenate(parts)
        out, info = oracle.histogram_sort(S, list(sizes), list(sizes), key)
        assert info['rc'] == 0
        S2 = S.copy()
        S2.sort()
        assert numpy.array_equal(out['vkey'], S2['vkey'])


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("nkeys", [None, 1000, 7])
def test_histogram_sort_matches_reference_mpsort_omp(nkeys):
    """mpsort_omp (mpsort-omp.c) runs the same algorithm with T threads as
    ranks and an even split; its bytes equal the restatement's."""
    a = _recs(20000, 20, nkeys)
    key = oracle.key_for(a.dtype, 'key')
    T = 4
    n = len(a)
    cnt = [n * (t + 1) // T - n * t // T for t in range(T)]
    out, info = oracle.histogram_sort(a, cnt, cnt, key)
    ref = oracle.ref_mpsort_omp(a, key, nthreads=T)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))


@pytest.mark.skipif(not oracle.mpi_available(), reason="reference MPI driver or mpiexec not available")
@pytest.mark.parametrize("nin,nout,nkeys", [
    ([2500, 2500, 2500, 2500], [2500, 2500, 2500, 2500], None),
    ([0, 4000, 0, 6000], [2000, 2000, 0, 6000], 1000),
    ([3000, 1000, 4000, 2000], [1000, 4000, 2000, 3000], 7),
])
def test_histogram_sort_matches_reference_mpsort_mpi(nin, nout, nkeys):
    a = _recs(sum(nin), 30, nkeys)
    key = oracle.key_for(a.dtype, 'key')
    out, info = oracle.histogram_sort(a, nin, nout, key)
    ref, _ = oracle.ref_mpi_sort(a, nin, nout, key)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))


@pytest.mark.skipif(not oracle.mpi_available(), reason="reference MPI driver or mpiexec not available")
def test_reference_mpsort_mpi_signed_and_multiword():
    dt = numpy.dtype([('v', 'u4'), ('k', ('i4', 2)), ('pad', 'u4')])
    g = rng(7)
    a = numpy.zeros(4000, dt)
    a['k'] = g.integers(-50, 50, (4000, 2))
    a['v'] = numpy.arange(4000)
    key = oracle.key_for(dt, 'k')
    nin = [1000, 1000, 1000, 1000]
    out, info = oracle.histogram_sort(a, nin, nin, key)
    ref, _ = oracle.ref_mpi_sort(a, nin, nin, key)
    assert numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8))


# ---- the learned path's root network (oracle_mlp_grad) ---------------------
def test_mlp_oracle_gradient_matches_finite_differences():
    """oracle_mlp_grad is pinned by central differences of oracle_mlp_loss;
    a dropped chain-rule factor (softplus'/tanh') fails this."""
    g = rng(0)
    raw = g.normal(0, 1, oracle.MLP_PARAMS)
    x = g.random(64).astype('f4')
    y = numpy.sort(g.random(64)).astype('f4')
    loss, grad = oracle.mlp_grad(raw, x, y)
    assert abs(loss - oracle.mlp_loss(raw, x, y)) < 1e-12
    for i in list(range(0, oracle.MLP_PARAMS, 37)) + [oracle.MLP_PARAMS - 1]:
        h = 1e-6
        a = raw.copy(); a[i] += h
        b = raw.copy(); b[i] -= h
        fd = (oracle.mlp_loss(a, x, y) - oracle.mlp_loss(b, x, y)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_mlp_oracle_is_monotone():
    g = rng(1)
    raw = g.normal(0, 2, oracle.MLP_PARAMS)
    xs = numpy.linspace(0, 1, 2001)
    ys = [oracle.mlp_forward(raw, float(v)) for v in xs]
    assert all(b >= a - 1e-12 for a, b in zip(ys, ys[1:]))


# ---- the reference's Issue7 data (test_mpsort.py:354-556) -----------------
def _issue7():
    import os
    z = numpy.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "issue7.npz"))
    return [int(c) for c in z['counts']], z['records'], z['expected'], z['words']


def test_issue7_fixture_is_the_reference_test_data():
    """The fixture holds the 12 rank chunks of Issue7B64 and the records
    test_issue7 builds from them: radix word 1 = column 0, radix word 0 =
    column 5 (column 4 << 32 is 0 in u4 arithmetic)."""
    counts, recs, expected, words = _issue7()
    assert counts == [0, 0, 2, 7, 1, 0, 4, 0, 0, 6, 14, 5]
    assert words.shape == (39, 10) and recs.dtype.itemsize == 40
    assert numpy.array_equal(recs['radix'][:, 1], words[:, 0])
    assert numpy.array_equal(recs['radix'][:, 0], words[:, 5])
    assert not recs['ext'].any()
    # the answer is the stable order of (word 1, word 0)
    order = sorted(range(len(recs)), key=lambda i: (int(recs['radix'][i, 1]), int(recs['radix'][i, 0])))
    assert numpy.array_equal(expected, recs[order])


def test_issue7_oracle_matches_fixture():
    counts, recs, expected, _ = _issue7()
    key = oracle.key_for(recs.dtype, 'radix')
    out, info = oracle.histogram_sort(recs, counts, counts, key)
    assert info['rc'] == 0
    assert numpy.array_equal(out, expected)


def test_issue7_reference_matches_fixture():
    if not oracle.ref_available():
        pytest.skip("compiled reference not built (oracle/_ref)")
    counts, recs, expected, _ = _issue7()
    key = oracle.key_for(recs.dtype, 'radix')
    assert numpy.array_equal(oracle.ref_radix_sort(recs, key), expected)
    if oracle.mpi_available():
        out, _ = oracle.ref_mpi_sort(recs, counts, counts, key)
        assert numpy.array_equal(out, expected)
