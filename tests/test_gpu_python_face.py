"""The Python face (mpsort.sort / binding.sort / permute / take / VirtualComm)
on the GPU, case for case after MP-sort's own tests
(mpsort/tests/test_mpsort.py:29-351, 531-556).

The reference runs each case on MPI ranks; here one process holds the whole
array and a VirtualComm splits it into virtual ranks with the same per-rank
input and output counts (split(), heal() and adjustsize() of the reference
become the in_sizes / out_sizes below).  Each result is checked against
numpy's stable sort or lexsort of the same data, and the Issue7 case against
the committed fixture decoded from the reference's own data."""
import os
from itertools import product

import numpy
import pytest
from numpy.testing import assert_array_equal

import mpsort
from mpsort import VirtualComm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

TUNINGS = [
    [],
    ['DISABLE_SPARSE_ALLTOALLV'],
    ['REQUIRE_SPARSE_ALLTOALLV'],
    ['REQUIRE_GATHER_SORT'],
    ['DISABLE_GATHER_SORT'],
]
RANKS = 4


def even_sizes(n, P=RANKS):
    """numpy.array_split sizes (test_mpsort.py:8-15)"""
    return [len(x) for x in numpy.array_split(numpy.arange(n), P)]


def adjusted(sizes):
    """adjustsize() of test_mpsort.py:22-27: rank r gets size + 1 - 2 * (r % 2)
    (an odd rank count leaves rank 0 alone)"""
    out = [s + 1 - 2 * (r % 2) for r, s in enumerate(sizes)]
    if len(sizes) % 2 == 1:
        out[0] = sizes[0]
    return out


def comm_for(n, out_sizes=None, in_sizes=None):
    ins = in_sizes if in_sizes is not None else even_sizes(n)
    return VirtualComm(in_sizes=ins, out_sizes=out_sizes if out_sizes is not None else ins)


def rng(seed):
    return numpy.random.default_rng(seed)


@pytest.mark.parametrize("dtype", ['i4', 'i8', 'u4', 'u8'])
def test_sort_scalar_dtypes(dtype):
    """test_sort_i4 / i8 / u8 / u4 (test_mpsort.py:29-83)"""
    g = rng(1)
    if dtype[0] == 'i':
        s = (g.random(1000) * 1000 - 400).astype(dtype)
    elif dtype == 'u4':
        s = g.integers(0, 2 ** 32, 1000, dtype=numpy.uint64).astype('u4')
    else:
        s = g.integers(0, 2 ** 64, 1000, dtype=numpy.uint64)
    local = s.copy()
    mpsort.sort(local, orderby=None, out=None, comm=comm_for(len(s)))
    assert_array_equal(local, numpy.sort(s, kind='stable'))


@pytest.mark.parametrize("dtype", ['i4', 'i8'])
def test_sort_signed_extremes(dtype):
    info = numpy.iinfo(dtype)
    s = numpy.array([info.max, -1, 0, info.min, 1, info.min + 1, info.max - 1, -2] * 40, dtype=dtype)
    local = s.copy()
    mpsort.sort(local, comm=comm_for(len(s)))
    assert_array_equal(local, numpy.sort(s))


@pytest.mark.parametrize("tuning", TUNINGS)
def test_sort_tunings(tuning):
    """test_mpsort.py:85-107"""
    s = (rng(2).random(1000) * 1000).astype('i4')
    local = s.copy()
    mpsort.sort(local, orderby=None, out=None, comm=comm_for(len(s)), tuning=tuning)
    assert_array_equal(local, numpy.sort(s))


def test_sort_inplace_by_array():
    """test_sort_inplace (test_mpsort.py:110-123): orderby is the array itself"""
    s = (rng(3).random(1000) * 1000).astype('i4')
    local = s.copy()
    mpsort.sort(local, local, out=None, comm=comm_for(len(s)))
    assert_array_equal(local, numpy.sort(s))


def test_sort_mismatched_zeros():
    """test_mpsort.py:125-140: input ranks [0, 400, 0, 600], output ranks
    [200, 200, 0, 600], REQUIRE_GATHER_SORT, out pre-filled with noise"""
    g = rng(4)
    s = (g.random(1000) * 1000).astype('i4')
    local = s.copy()
    res = (g.random(1000) * 1000).astype('i4')
    comm = VirtualComm(in_sizes=[0, 400, 0, 600], out_sizes=[200, 200, 0, 600])
    mpsort.sort(local, local, out=res, comm=comm, tuning=['REQUIRE_GATHER_SORT'])
    assert_array_equal(res, numpy.sort(s))
    assert_array_equal(local, s)


def test_sort_outplace_skewed_output():
    """test_sort_outplace (test_mpsort.py:142-156): out ranks of size
    adjustsize(local)"""
    s = (rng(5).random(1000) * 1000).astype('i4')
    local = s.copy()
    ins = even_sizes(len(s))
    res = numpy.zeros(len(s), dtype=local.dtype)
    mpsort.sort(local, local, out=res, comm=VirtualComm(in_sizes=ins, out_sizes=adjusted(ins)))
    assert_array_equal(res, numpy.sort(s))


def test_sort_flatiter():
    """test_mpsort.py:158-174: source, orderby and out as flat iterators"""
    s = (rng(6).random(1000) * 1000).astype('i4')
    local = s.copy()
    ins = even_sizes(len(s))
    res = numpy.zeros(len(s), dtype=local.dtype)
    mpsort.sort(local.flat, local.flat, out=res.flat, comm=VirtualComm(in_sizes=ins, out_sizes=adjusted(ins)))
    assert_array_equal(res, numpy.sort(s))


def test_sort_flatiter_of_2d_array():
    """a flat iterator stands for every scalar of its array"""
    s = (rng(7).random((50, 4)) * 1000).astype('i8')
    res = numpy.zeros((50, 4), dtype='i8')
    mpsort.sort(s.flat, out=res.flat, comm=VirtualComm(vranks=3))
    assert_array_equal(res.reshape(-1), numpy.sort(s.reshape(-1)))


def test_sort_struct_by_field():
    """test_sort_struct (test_mpsort.py:176-199)"""
    s = numpy.empty(10, dtype=[('value', 'i8'), ('key', 'i8')])
    g = rng(1234)
    s['value'] = (g.random(10) * 1000 - 400).astype('i4')
    s['key'] = s['value']
    backup = s.copy()
    res = numpy.zeros_like(s)
    mpsort.sort(s, 'key', out=res, comm=comm_for(len(s)))
    order = numpy.argsort(backup['key'], kind='stable')
    assert_array_equal(res, backup[order])


def test_sort_struct_by_field_stable_with_duplicates():
    g = rng(12)
    s = numpy.empty(5000, dtype=[('value', 'i8'), ('key', 'i8')])
    s['key'] = g.integers(-5, 5, len(s))
    s['value'] = numpy.arange(len(s))
    res = numpy.zeros_like(s)
    mpsort.sort(s, 'key', out=res, comm=comm_for(len(s)))
    assert_array_equal(res, s[numpy.argsort(s['key'], kind='stable')])


def test_sort_struct_vector_key():
    """test_sort_struct_vector (test_mpsort.py:201-226): a 2-word key field,
    the later word more significant"""
    s = numpy.empty(10, dtype=[('value', 'i8'), ('key', 'i8'), ('vkey', ('i8', 2))])
    s['value'] = (rng(8).random(len(s)) * 1000).astype('i4')
    s['key'] = s['value']
    s['vkey'][:, 0] = s['value']
    s['vkey'][:, 1] = s['value']
    res = numpy.zeros_like(s)
    mpsort.sort(s, 'vkey', out=res, comm=comm_for(len(s)))
    order = numpy.argsort(s['key'], kind='stable')
    assert_array_equal(res['value'], s['value'][order])


def test_sort_struct_vector_key_word_order():
    """distinct words: lexsort with the last word as the primary key"""
    g = rng(9)
    s = numpy.empty(3000, dtype=[('value', 'i8'), ('vkey', ('i8', 2))])
    s['vkey'][:, 0] = g.integers(-50, 50, len(s))
    s['vkey'][:, 1] = g.integers(-3, 3, len(s))
    s['value'] = numpy.arange(len(s))
    res = numpy.zeros_like(s)
    mpsort.sort(s, 'vkey', out=res, comm=comm_for(len(s)))
    order = numpy.lexsort(s['vkey'].T)
    assert_array_equal(res, s[order])


def test_sort_vector_key_array():
    """test_sort_vector (test_mpsort.py:228-250): orderby is an (n, 2) array"""
    s = numpy.empty(10, dtype=[('value', 'i8')])
    s['value'] = (rng(10).random(len(s)) * 1000).astype('i4')
    k = numpy.empty((len(s), 2), 'i8')
    k[:, 0] = s['value']
    k[:, 1] = s['value']
    res = numpy.zeros_like(s)
    mpsort.sort(s, k, out=res, comm=comm_for(len(s)))
    assert_array_equal(res['value'], numpy.sort(s['value']))


def test_sort_by_key_array_is_stable():
    g = rng(11)
    s = numpy.arange(20000, dtype='i8')
    k = g.integers(0, 7, len(s)).astype('u4')
    res = numpy.zeros_like(s)
    mpsort.sort(s, k, out=res, comm=VirtualComm(vranks=16))
    assert_array_equal(res, s[numpy.argsort(k, kind='stable')])


def _permute_case():
    s = numpy.arange(10)
    i = numpy.arange(9, -1, -1)
    return s, i


def test_permute():
    """test_permute (test_mpsort.py:253-265)"""
    s, i = _permute_case()
    res = mpsort.permute(s, i, comm_for(10))
    assert res.size == i.size
    assert_array_equal(res, s[i])


def test_permute_out():
    """test_permute_out (test_mpsort.py:267-279)"""
    s, i = _permute_case()
    res = numpy.empty(s.size, s.dtype)
    mpsort.permute(s, i, comm_for(10), out=res)
    assert_array_equal(res, s[i])


def test_permute_size_mismatch():
    with pytest.raises(ValueError, match="Global size"):
        mpsort.permute(numpy.arange(10), numpy.arange(9), VirtualComm())


def test_take():
    """test_take (test_mpsort.py:281-293)"""
    s, i = _permute_case()
    res = mpsort.take(s, i, comm_for(10))
    assert res.size == i.size
    assert_array_equal(res, s[i])


def test_take_out():
    """test_take_out (test_mpsort.py:295-307)"""
    s, i = _permute_case()
    res = numpy.empty(s.size, s.dtype)
    mpsort.take(s, i, comm_for(10), out=res)
    assert_array_equal(res, s[i])


def test_take_repeated_indices_and_records():
    """argindex need be neither a permutation nor unique; records with
    subarrays come along whole"""
    g = rng(13)
    s = numpy.zeros(5000, dtype=[('id', 'i8'), ('pos', 'f4', 3)])
    s['id'] = numpy.arange(len(s))
    s['pos'] = g.random((len(s), 3))
    i = g.integers(0, len(s), 12345)
    res = mpsort.take(s, i, VirtualComm(vranks=8))
    assert_array_equal(res, s[i])


def test_permute_large_random():
    g = rng(14)
    n = 1 << 20
    s = g.integers(0, 2 ** 62, n)
    i = g.permutation(n)
    res = mpsort.permute(s, i, VirtualComm(vranks=64))
    assert_array_equal(res, s[i])


@pytest.mark.parametrize("sizes, tuning", list(product(product(*([[0, 1, 2]] * 4)), TUNINGS)))
def test_few_items(sizes, tuning):
    """test_few_items (test_mpsort.py:330-351): rank r holds range(sizes[r])
    as a 3-word u8 key, output on the same partition"""
    parts = []
    for r in range(4):
        s = numpy.empty(sizes[r], dtype=[('vkey', ('u8', 3)), ('vector', ('u4', 3))])
        s['vkey'] = numpy.arange(sizes[r], dtype='u8')[:, None]
        s['vector'] = 1
        parts.append(s)
    S = numpy.concatenate(parts)
    want = S.copy()
    want.sort()
    r = numpy.empty(len(S), dtype=S.dtype)
    mpsort.sort(S, 'vkey', out=r, comm=VirtualComm(in_sizes=list(sizes), out_sizes=list(sizes)), tuning=tuning)
    if len(S):
        assert_array_equal(r['vkey'], want['vkey'])


def test_binding_sort_directly():
    """binding.sort (binding.pyx:123-209) with a VirtualComm"""
    g = rng(15)
    s = numpy.zeros(3000, dtype=[('key', 'u8'), ('val', 'i8')])
    s['key'] = g.integers(0, 100, len(s))
    s['val'] = numpy.arange(len(s))
    out = numpy.empty_like(s)
    mpsort.binding.sort(s, 'key', out, comm=VirtualComm(vranks=5))
    assert_array_equal(out, s[numpy.argsort(s['key'], kind='stable')])
    inplace = s.copy()
    mpsort.binding.sort(inplace, 'key', comm=VirtualComm(vranks=5))
    assert_array_equal(inplace, out)


@pytest.mark.parametrize("tuning", TUNINGS)
def test_issue7(tuning):
    """test_issue7 (test_mpsort.py:531-556) on the reference's own data
    (tests/golden/issue7.npz, decoded from Issue7B64): 40-byte records with a
    16-byte key over 12 ranks"""
    z = numpy.load(os.path.join(HERE, "golden", "issue7.npz"))
    counts = [int(c) for c in z['counts']]
    s = z['records'].copy()
    r = numpy.empty_like(s)
    mpsort.sort(s, orderby='radix', out=r, comm=VirtualComm(in_sizes=counts, out_sizes=counts), tuning=tuning)
    assert_array_equal(r, z['expected'])


def test_issue7_shape_at_scale():
    """the Issue7 record shape (40 B, 2 x u8 key) at 2^20 records over 12
    ranks of uneven size, against lexsort"""
    g = rng(16)
    n = 1 << 20
    s = numpy.zeros(n, dtype=[('radix', ('u8', 2)), ('ext', ('u8', 3))])
    s['radix'][:, 0] = g.integers(0, 2 ** 40, n)
    s['radix'][:, 1] = g.integers(0, 300, n)
    s['ext'][:, 0] = numpy.arange(n)
    w = g.random(12) + 0.05
    sizes = numpy.floor(w / w.sum() * n).astype(int)
    sizes[-1] += n - sizes.sum()
    r = numpy.empty_like(s)
    mpsort.sort(s, orderby='radix', out=r, comm=VirtualComm(in_sizes=list(sizes), out_sizes=list(sizes[::-1])))
    assert_array_equal(r, s[numpy.lexsort(s['radix'].T)])
