"""The exchange's two run tables (MPIU_Alltoallv mp-mpiu.c:125-236): the
dense one walks every (sender, receiver) slot, the sparse one only the
non-empty runs.  Both give the oracle's bytes and counts; the choice follows
the call's flags, then the recorded options (REQUIRE over DISABLE,
mpsort-mpi.c:576-583), then the reference's AUTO vote (dense when some sender
has more than 128 non-empty receivers, mp-mpiu.c:125-129); sparse falls back
to dense when the non-empty runs exceed 64 per rank on average.  With no
table chosen and at most 64 ranks there is no exchange at all: the SecondSort
gathers every receiver bucket from the senders' sorted segments."""
import sys

import numpy
import pytest

import mpsort
from gpu_util import even
from mpsort import _lib, workloads
from mpsort._lib import lib
from test_gpu_histsort import check_parity, rec16

pytestmark = pytest.mark.gpu

DENSE, SPARSE = _lib.EXCHANGE_DENSE, _lib.EXCHANGE_SPARSE


def note(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nearly_sorted16(n, seed, frac=0.001):
    """sorted distinct keys with round(frac * n) records moved to other
    chosen positions (the snapshot-B shape: B derived from sorted A)"""
    g = numpy.random.default_rng(seed)
    a = numpy.zeros(n, workloads.REC16)
    a['key'] = numpy.sort(g.integers(0, 2 ** 64, n, dtype=numpy.uint64))
    m = max(2, int(round(frac * n)))
    pos = numpy.sort(g.choice(n, m, replace=False))
    a['key'][pos] = a['key'][numpy.roll(pos, 1)]
    a['idx'] = numpy.arange(n)
    return a


def expected(policy, P, send):
    """(sparse?, runs walked, merged?) for the oracle's send matrix"""
    nnz = int(numpy.count_nonzero(send))
    cap = min(P * P, 64 * P)
    peers = int((numpy.asarray(send) > 0).sum(axis=1).max())
    if policy == "auto" and P <= 64:
        return False, 0, True
    if policy == "dense":
        sparse = False
    elif policy == "sparse":
        sparse = nnz <= cap
    else:
        sparse = peers <= 128 and nnz <= cap
    return sparse, (nnz if sparse else P * P), False


@pytest.mark.parametrize("shape", ["random", "nearly_sorted"])
@pytest.mark.parametrize("P", [2, 3, 64, 130, 300])
def test_dense_and_sparse_tables_match_the_oracle(P, shape):
    n = 4000 + 53 * P
    a = rec16(n, 11 + P) if shape == "random" else nearly_sorted16(n, 11 + P)
    import oracle
    key = oracle.key_for(a.dtype, 'key')
    ref, info = oracle.histogram_sort(a, even(n, P), even(n, P), key)
    outs = []
    for policy, flags in (("dense", DENSE), ("sparse", SPARSE), ("auto", 0), ("vote", _lib.EXCHANGE)):
        note("exchange P=%d shape=%s policy=%s" % (P, shape, policy))
        out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=flags, device=True)
        sp, runs, merged = expected(policy, P, info['send'])
        assert (st.exchange_sparse, st.exchange_runs, st.second_merged) == (int(sp), runs, int(merged)), policy
        outs.append(out)
    assert all(numpy.array_equal(outs[0], o) for o in outs[1:])


def test_nearly_sorted_input_walks_few_runs():
    """snapshot-B shape on 64 ranks: a receiver gets its own sender's run and
    a few displaced records, so the sparse table is a small fraction of the
    4096 slots"""
    n, P = 1 << 16, 64
    a = nearly_sorted16(n, 5)
    out, st = check_parity(a, even(n, P), even(n, P), 'key', flags=_lib.EXCHANGE, device=True)
    assert st.exchange_sparse == 1 and st.second_merged == 0
    assert P <= st.exchange_runs < P * P // 8


@pytest.mark.parametrize("opts, sparse", [
    (mpsort.binding._TUNING['REQUIRE_SPARSE_ALLTOALLV'], 1),
    (mpsort.binding._TUNING['DISABLE_SPARSE_ALLTOALLV'], 0),
    (mpsort.binding._TUNING['REQUIRE_SPARSE_ALLTOALLV'] | mpsort.binding._TUNING['DISABLE_SPARSE_ALLTOALLV'], 1),
])
def test_recorded_options_choose_the_table(opts, sparse):
    n, P = 20000, 8
    a = rec16(n, 3)
    lib.mpsort_mpi_unset_options(-1)
    lib.mpsort_mpi_set_options(opts)
    try:
        out, st = check_parity(a, even(n, P), even(n, P), 'key', device=True)
    finally:
        lib.mpsort_mpi_unset_options(-1)
    assert st.exchange_sparse == sparse


@pytest.mark.parametrize("tuning, sparse, merged", [([], 0, 1), (['DISABLE_SPARSE_ALLTOALLV'], 0, 0),
                                                    (['REQUIRE_SPARSE_ALLTOALLV'], 1, 0)])
def test_python_tuning_names(tuning, sparse, merged):
    """test_mpsort.py:85-107: the tunings sort the same; here they also pick
    the run table (none named: no exchange, the SecondSort gathers from the
    senders' sorted segments)"""
    g = numpy.random.default_rng(9)
    a = g.integers(-2 ** 62, 2 ** 62, 30000).astype('i8')
    want = numpy.sort(a, kind='stable')
    try:
        mpsort.sort(a, orderby=None, comm=mpsort.VirtualComm(vranks=7), tuning=tuning)
        st = mpsort.last_stats()
    finally:
        # the tuning names stay recorded after the call, as in the reference
        # binding (binding.pyx:195-204): clear them for the later tests
        lib.mpsort_mpi_unset_options(-1)
    assert numpy.array_equal(a, want)
    assert (st['exchange_sparse'], st['second_merged']) == (sparse, merged)
