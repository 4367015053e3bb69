"""ctypes wrappers of the oracle (liboracle.so) and of the compiled reference
(_ref/).  Test infrastructure only."""
import ctypes
import os
import shutil
import subprocess
import tempfile
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmpsort_ref.so")
REF_DEFAULT_SO = os.path.join(HERE, "_ref", "libmpsort_ref_default.so")
REF_MPI = os.path.join(HERE, "_ref", "ref_mpi")

__all__ = ["OKey", "key_for", "radix", "compar", "bisect", "radix_sort", "histogram_sort",
           "closed_form", "stable_order", "ref_available", "ref_radix_sort", "ref_mpsort_omp",
           "mpi_available", "ref_mpi_sort", "ORACLE_SO", "REF_SO", "REF_MPI", "omp_max_threads"]


class OKey(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_int64), ("width", ctypes.c_int32),
                ("is_signed", ctypes.c_int32), ("nwords", ctypes.c_int32), ("pad", ctypes.c_int32)]


class OResult(ctypes.Structure):
    _fields_ = [("P", ctypes.c_void_p), ("Pmin", ctypes.c_void_p), ("Pmax", ctypes.c_void_p),
                ("C", ctypes.c_void_p), ("myCLT", ctypes.c_void_p), ("myCLE", ctypes.c_void_p),
                ("myC", ctypes.c_void_p), ("send", ctypes.c_void_p), ("iterations", ctypes.c_int64),
                ("checksum_in", ctypes.c_int64), ("checksum_out", ctypes.c_int64)]


_lib = None


def _oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError("oracle/liboracle.so missing: run `make -C oracle oracle`")
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.oracle_compar.restype = ctypes.c_int
        _lib.oracle_histogram_sort.restype = ctypes.c_int
    return _lib


def key_for(dtype, field=None):
    """Key descriptor of dtype[field] (binding.pyx:50-79 semantics)."""
    dtype = numpy.dtype(dtype)
    if field is not None:
        kd, off = dtype.fields[field][:2]
    else:
        kd, off = dtype, 0
    nw = 1 if len(kd.shape) == 0 else kd.shape[0]
    b = kd.base
    table = {numpy.dtype('u8'): (8, 0), numpy.dtype('i8'): (8, 1),
             numpy.dtype('u4'): (4, 0), numpy.dtype('i4'): (4, 1)}
    if b not in table:
        raise TypeError("key type %s not supported" % b)
    w, s = table[b]
    return OKey(offset=int(off), width=w, is_signed=s, nwords=int(nw), pad=0)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def radix(records, key):
    """Radix bytes (n x 8*nwords) of every record: binding.pyx:81-121."""
    records = numpy.ascontiguousarray(records)
    n = len(records)
    rs = 8 * key.nwords
    out = numpy.zeros((n, rs), dtype=numpy.uint8)
    _oracle().oracle_radix_batch(_ptr(records), ctypes.c_size_t(n),
                                 ctypes.c_size_t(records.dtype.itemsize), ctypes.byref(key), _ptr(out))
    return out


def compar(a, b):
    a = numpy.ascontiguousarray(a, dtype=numpy.uint8)
    b = numpy.ascontiguousarray(b, dtype=numpy.uint8)
    return _oracle().oracle_compar(_ptr(a), _ptr(b), ctypes.c_size_t(len(a)))


def bisect(a, b):
    a = numpy.ascontiguousarray(a, dtype=numpy.uint8)
    b = numpy.ascontiguousarray(b, dtype=numpy.uint8)
    m = numpy.zeros_like(a)
    _oracle().oracle_bisect(_ptr(m), _ptr(a), _ptr(b), ctypes.c_size_t(len(a)))
    return m


def radix_sort(records, key):
    """Stable sort by radix (radixsort.c:35-44); returns a sorted copy."""
    out = numpy.array(records, copy=True)
    _oracle().oracle_radix_sort(_ptr(out), ctypes.c_size_t(len(out)),
                                ctypes.c_size_t(out.dtype.itemsize), ctypes.byref(key))
    return out


def histogram_sort(records, nin, nout, key=None, radix_bytes=None):
    """Restated mpsort_mpi_newarray over len(nin) ranks held in one process.

    Returns (out, info) where info holds P (splitters, (P-1, rsize) bytes),
    Pmin, Pmax, C, myCLT, myCLE, myC (P x (P+1)), send (P x P), iterations,
    checksum_in / checksum_out and rc."""
    records = numpy.ascontiguousarray(records)
    P = len(nin)
    nin = numpy.asarray(nin, dtype=numpy.int64)
    nout = numpy.asarray(nout, dtype=numpy.int64)
    if radix_bytes is not None:
        radix_bytes = numpy.ascontiguousarray(radix_bytes, dtype=numpy.uint8)
        rs = radix_bytes.shape[1] if radix_bytes.ndim == 2 else 0
    else:
        rs = 8 * key.nwords
    out = numpy.zeros_like(records) if len(records) else records.copy()
    Ps = numpy.zeros((max(P - 1, 1), max(rs, 1)), dtype=numpy.uint8)
    pmin = numpy.zeros(max(rs, 1), dtype=numpy.uint8)
    pmax = numpy.zeros(max(rs, 1), dtype=numpy.uint8)
    C = numpy.zeros(P + 1, dtype=numpy.int64)
    clt = numpy.zeros((P, P + 1), dtype=numpy.int64)
    cle = numpy.zeros((P, P + 1), dtype=numpy.int64)
    myc = numpy.zeros((P, P + 1), dtype=numpy.int64)
    send = numpy.zeros((P, P), dtype=numpy.int64)
    res = OResult(P=_ptr(Ps).value, Pmin=_ptr(pmin).value, Pmax=_ptr(pmax).value, C=_ptr(C).value,
                  myCLT=_ptr(clt).value, myCLE=_ptr(cle).value, myC=_ptr(myc).value,
                  send=_ptr(send).value)
    rec = records if len(records) else numpy.zeros(1, dtype=records.dtype)
    o = out if len(out) else numpy.zeros(1, dtype=records.dtype)
    rc = _oracle().oracle_histogram_sort(
        _ptr(rec), _ptr(nin), ctypes.c_int(P), _ptr(o), _ptr(nout),
        ctypes.c_size_t(records.dtype.itemsize), ctypes.byref(key) if key is not None else None,
        _ptr(radix_bytes) if radix_bytes is not None else None, ctypes.c_size_t(rs), ctypes.byref(res))
    info = dict(rc=rc, P=Ps[:P - 1].copy(), Pmin=pmin[:rs].copy(), Pmax=pmax[:rs].copy(), C=C,
                myCLT=clt, myCLE=cle, myC=myc, send=send, iterations=res.iterations,
                checksum_in=res.checksum_in, checksum_out=res.checksum_out)
    return out, info


def stable_order(radix_bytes):
    """Stable argsort of radix rows read as little-endian unsigned integers,
    computed with numpy (an implementation independent of liboracle)."""
    r = numpy.ascontiguousarray(radix_bytes, dtype=numpy.uint8)
    n, rs = r.shape
    order = numpy.arange(n)
    # LSD over bytes with numpy's stable sort: least significant byte first
    for b in range(rs):
        order = order[numpy.argsort(r[order, b], kind='stable')]
    return order


def closed_form(radix_bytes, nin, nout):
    """Brute-force values every histogram sort must produce: the globally
    stable order, P_i = k[C[i+1]-1] (Pmin for C[i+1] == 0), per-rank counts of
    radix < / <= P_i, and the send matrix read off the stable order."""
    r = numpy.ascontiguousarray(radix_bytes, dtype=numpy.uint8)
    n, rs = r.shape
    P = len(nin)
    nin = numpy.asarray(nin, dtype=numpy.int64)
    nout = numpy.asarray(nout, dtype=numpy.int64)
    order = stable_order(r)
    as_int = [int.from_bytes(bytes(row), 'little') for row in r]
    k = [as_int[i] for i in order]
    C = numpy.concatenate([[0], numpy.cumsum(nout)]).astype(numpy.int64)
    start = numpy.concatenate([[0], numpy.cumsum(nin)]).astype(numpy.int64)
    nonempty = [t for t in range(P) if nin[t] > 0]
    pmin = min(as_int) if n else 0
    splitters = []
    for i in range(P - 1):
        c = int(C[i + 1])
        splitters.append(k[c - 1] if c >= 1 else pmin)
    clt = numpy.zeros((P, P + 1), dtype=numpy.int64)
    cle = numpy.zeros((P, P + 1), dtype=numpy.int64)
    for j in range(P):
        mine = sorted(as_int[start[j]:start[j + 1]])
        for i, s in enumerate(splitters):
            clt[j, i + 1] = sum(1 for x in mine if x < s)
            cle[j, i + 1] = sum(1 for x in mine if x <= s)
        clt[j, P] = cle[j, P] = len(mine)
    # send[j, r] = records of rank j whose position in the stable order lies
    # in receiver r's output range
    src_rank = numpy.searchsorted(start[1:], order, side='right')
    pos_rank = numpy.searchsorted(C[1:], numpy.arange(n), side='right')
    send = numpy.zeros((P, P), dtype=numpy.int64)
    numpy.add.at(send, (src_rank, pos_rank), 1)
    return dict(order=order, splitters=splitters, C=C, myCLT=clt, myCLE=cle, send=send,
                nonempty=nonempty)


# ---- the compiled reference -------------------------------------------------
_ref = {}


def ref_available(default_flags=False):
    return os.path.exists(REF_DEFAULT_SO if default_flags else REF_SO)


def _reflib(default_flags=False):
    path = REF_DEFAULT_SO if default_flags else REF_SO
    if path not in _ref:
        if not os.path.exists(path):
            raise ImportError("%s missing: run `make -C oracle ref` where the reference tree exists" % path)
        l = ctypes.CDLL(path)
        l.ref_omp_max_threads.restype = ctypes.c_int
        _ref[path] = l
    return _ref[path]


def omp_max_threads():
    return _reflib().ref_omp_max_threads()


def ref_radix_sort(records, key, default_flags=False):
    out = numpy.array(records, copy=True)
    if len(out):
        _reflib(default_flags).ref_radix_sort(_ptr(out), ctypes.c_size_t(len(out)),
                                              ctypes.c_size_t(out.dtype.itemsize), ctypes.byref(key))
    return out


def ref_mpsort_omp(records, key, nthreads=0, default_flags=False, inplace=False, timed=False):
    """The compiled reference mpsort_omp; with timed=True also the seconds of
    the call itself (the copy of the input made before it is not counted)."""
    out = records if inplace else numpy.array(records, copy=True)
    t = 0.0
    if len(out):
        t0 = time.perf_counter()
        _reflib(default_flags).ref_mpsort_omp(_ptr(out), ctypes.c_size_t(len(out)),
                                              ctypes.c_size_t(out.dtype.itemsize), ctypes.byref(key),
                                              ctypes.c_int(nthreads))
        t = time.perf_counter() - t0
    return (out, t) if timed else out


def _mpiexec():
    for c in (shutil.which("mpiexec"), "/opt/conda/bin/mpiexec"):
        if c and os.path.exists(c):
            return c
    return None


def mpi_available():
    return os.path.exists(REF_MPI) and _mpiexec() is not None


def ref_mpi_sort(records, nin, nout, key, repeat=1, timeout=600):
    """Run the compiled reference mpsort_mpi_newarray on len(nin) MPI ranks
    (oversubscribed if needed); returns (sorted records, median seconds)."""
    records = numpy.ascontiguousarray(records)
    P = len(nin)
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        records.tofile(fin)
        cmd = [_mpiexec(), "-n", str(P), REF_MPI, fin, fout, str(records.dtype.itemsize),
               str(key.offset), str(key.width), str(key.is_signed), str(key.nwords), str(repeat), "0"]
        cmd += [str(int(x)) for x in nin] + [str(int(x)) for x in nout]
        env = dict(os.environ)
        env.setdefault("HYDRA_LAUNCHER", "fork")
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError("ref_mpi failed: %s" % p.stderr.decode(errors="replace")[-2000:])
        t = None
        for line in p.stdout.decode().splitlines():
            if line.startswith("median_s"):
                t = float(line.split()[1])
        out = numpy.fromfile(fout, dtype=records.dtype)
    return out, t


# ---- the learned path's root network ---------------------------------------
MLP_PARAMS = 1153


def _mlp_lib():
    l = _oracle()
    l.oracle_mlp_grad.restype = ctypes.c_double
    l.oracle_mlp_loss.restype = ctypes.c_double
    l.oracle_mlp_forward.restype = ctypes.c_double
    return l


def mlp_grad(raw, x, y):
    """Double-precision loss and gradient (oracle.c oracle_mlp_grad)."""
    raw = numpy.ascontiguousarray(raw, dtype=numpy.float64)
    x = numpy.ascontiguousarray(x, dtype=numpy.float32)
    y = numpy.ascontiguousarray(y, dtype=numpy.float32)
    g = numpy.zeros(MLP_PARAMS, dtype=numpy.float64)
    loss = _mlp_lib().oracle_mlp_grad(_ptr(raw), _ptr(x), _ptr(y), ctypes.c_int64(len(x)), _ptr(g))
    return loss, g


def mlp_loss(raw, x, y):
    raw = numpy.ascontiguousarray(raw, dtype=numpy.float64)
    x = numpy.ascontiguousarray(x, dtype=numpy.float32)
    y = numpy.ascontiguousarray(y, dtype=numpy.float32)
    return _mlp_lib().oracle_mlp_loss(_ptr(raw), _ptr(x), _ptr(y), ctypes.c_int64(len(x)))


def mlp_forward(raw, x):
    raw = numpy.ascontiguousarray(raw, dtype=numpy.float64)
    return _mlp_lib().oracle_mlp_forward(_ptr(raw), ctypes.c_double(x))


__all__ += ["mlp_grad", "mlp_loss", "mlp_forward", "MLP_PARAMS"]
