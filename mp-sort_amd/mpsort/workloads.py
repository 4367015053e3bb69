"""Seeded synthetic workloads (libmpsort_gen.so, csrc/gen.c).

Inputs are generated once on the host; the same bytes go to the GPU path and
to the CPU checkers.  Parameters of every workload are pinned here.
"""
import ctypes
import os

import numpy

from ._lib import GENPATH

_g = None

REC16 = numpy.dtype([('key', '<u8'), ('idx', '<u8')])
REC16_SIGNED = numpy.dtype([('key', '<u8'), ('idx', '<i8')])
PARTICLE32 = numpy.dtype([('id', '<u8'), ('pos', '<f4', 3), ('vel', '<f4', 3)])
REC8 = numpy.dtype([('key', '<i4'), ('value', '<f4')])
REC16I = numpy.dtype([('key', '<i8'), ('value', '<i8')])

# uniform uint64 keys of the 4 x 1,000,000 record workload: PCG-XSL-RR-128/64
PCG_STATE = (0x0123456789ABCDEF, 0xFEDCBA9876543210)
PCG_INC = (0x5851F42D4C957F2D, 0x14057B7EF767814F)

# clustered Morton field: 2^16 clumps, median radius 2^-8, log-normal sigma 1,
# 30% uniform background, 21 bits per axis
CLUSTER = dict(nclump=1 << 16, median_r=2.0 ** -8, sigma_ln=1.0, frac_uniform=0.3)


def _gen():
    global _g
    if _g is None:
        if not os.path.exists(GENPATH):
            raise ImportError("libmpsort_gen.so missing: run `make -C mp-sort_amd`")
        _g = ctypes.CDLL(GENPATH)
        _g.mpsort_morton63.restype = ctypes.c_uint64
        _g.mpsort_gen_mix64.restype = ctypes.c_uint64
        _g.mpsort_gen_mix64.argtypes = [ctypes.c_uint64]
        _g.mpsort_gen_segment_sizes.restype = ctypes.c_uint64
    return _g


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


U64 = ctypes.c_uint64
SZ = ctypes.c_size_t


def pcg64(n, state=PCG_STATE, inc=PCG_INC):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_pcg64(_p(out), SZ(n), U64(state[0]), U64(state[1]), U64(inc[0]), U64(inc[1]))
    return out


def uniform(n, seed):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_uniform(_p(out), SZ(n), U64(seed))
    return out


def records16(keys, base=0):
    """{u64 key, u64 global index} records."""
    keys = numpy.ascontiguousarray(keys, dtype=numpy.uint64)
    out = numpy.empty(len(keys), dtype=REC16)
    _gen().mpsort_gen_records16(_p(out), _p(keys), SZ(len(keys)), U64(base))
    return out


def clustered_records(n, seed=2, nclump=None, median_r=None, sigma_ln=None, frac_uniform=None,
                      field_seed=None):
    """{u64 morton, i64 index} records of the clustered Morton field.

    The clumps come from field_seed (default: seed) and the points from
    seed, so records of one field_seed and two seeds are two realisations of
    one field."""
    c = dict(CLUSTER)
    for k, v in (("nclump", nclump), ("median_r", median_r), ("sigma_ln", sigma_ln),
                 ("frac_uniform", frac_uniform)):
        if v is not None:
            c[k] = v
    out = numpy.empty(n, dtype=REC16_SIGNED)
    _gen().mpsort_gen_clustered_field(None, _p(out), SZ(n), U64(seed if field_seed is None else field_seed),
                                      U64(seed), ctypes.c_int(int(c["nclump"])), ctypes.c_double(c["median_r"]),
                                      ctypes.c_double(c["sigma_ln"]), ctypes.c_double(c["frac_uniform"]))
    return out


def clustered_keys(n, seed=2, nclump=None):
    out = numpy.empty(n, dtype=numpy.uint64)
    c = dict(CLUSTER)
    if nclump is not None:
        c["nclump"] = nclump
    _gen().mpsort_gen_clustered(_p(out), None, SZ(n), U64(seed), ctypes.c_int(c["nclump"]),
                                ctypes.c_double(c["median_r"]), ctypes.c_double(c["sigma_ln"]),
                                ctypes.c_double(c["frac_uniform"]))
    return out


def morton63(x, y, z):
    return int(_gen().mpsort_morton63(ctypes.c_uint32(x), ctypes.c_uint32(y), ctypes.c_uint32(z)))


def lognormal_mix(n, seed):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_lognormal_mix(_p(out), SZ(n), U64(seed))
    return out


def zipf(n, seed, s=1.1, nvalues=1 << 20):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_zipf(_p(out), SZ(n), U64(seed), ctypes.c_double(s), ctypes.c_uint32(nvalues))
    return out


def nearly_sorted(n, seed, frac=0.01):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_nearly_sorted(_p(out), SZ(n), U64(seed), SZ(int(n * frac / 2)))
    return out


def permutation(n, seed):
    out = numpy.empty(n, dtype=numpy.uint64)
    _gen().mpsort_gen_permutation(_p(out), SZ(n), U64(seed))
    return out


def segment_sizes(nseg, seed, lo=1e3, hi=1e7, total=1 << 30):
    out = numpy.empty(nseg, dtype=numpy.uint64)
    _gen().mpsort_gen_segment_sizes(_p(out), SZ(nseg), U64(seed), ctypes.c_double(lo),
                                    ctypes.c_double(hi), U64(total))
    return out


def mix64(z):
    return int(_gen().mpsort_gen_mix64(z))


def snapshot_pair(n, seed, displaced=0.001):
    """Two snapshots of n particles ({u64 id, f32 pos[3], f32 vel[3]}, 32 B):
    in A the ids are a seeded permutation of 0..n-1 and pos/vel uniform in
    [0, 1); B is A sorted by id with round(displaced * n) records moved to
    other chosen positions (a cyclic shift among those positions): the next
    snapshot of a code that keeps its particles in id order, whose exchange
    is sparse (each rank keeps most of its records)."""
    a = numpy.empty(n, PARTICLE32)
    a['id'] = permutation(n, seed)
    g = numpy.random.default_rng(seed)
    a['pos'] = g.random((n, 3), dtype=numpy.float32)
    a['vel'] = g.random((n, 3), dtype=numpy.float32)
    inv = numpy.empty(n, numpy.int64)
    inv[a['id'].astype(numpy.int64)] = numpy.arange(n, dtype=numpy.int64)
    b = a[inv]  # sorted by id: the ids are a permutation of 0..n-1
    del inv
    m = int(round(displaced * n))
    if m > 1:
        pos = numpy.sort(g.choice(n, m, replace=False))
        b[pos] = b[numpy.roll(pos, 1)]
    return a, b
