"""Learned-splitter path: a key-CDF model trained on the GPU from the output
of one plain sort, then used by ``sort_learned`` (bit-identical output)."""
import ctypes

import numpy

from . import _lib
from ._lib import KeyDesc, ModelInfo, check, lib
from .binding import _key_desc


class CDFModel(object):
    """Monotone two-level key-CDF model (MLP root + linear leaves) held on the
    GPU.  Train it with :func:`train`."""

    def __init__(self, handle, key):
        self._h = handle
        self.key = key

    def info(self):
        i = ModelInfo()
        check(lib.mpsort_gpu_model_get_info(self._h, ctypes.byref(i)))
        return i.as_dict()

    def eval(self, keys):
        """F(key) for order-preserving uint64 keys."""
        keys = numpy.ascontiguousarray(keys, dtype=numpy.uint64)
        out = numpy.empty(len(keys), dtype=numpy.float32)
        check(lib.mpsort_gpu_model_eval(self._h, keys.ctypes.data, len(keys), out.ctypes.data))
        return out

    def params(self):
        raw = numpy.empty(_lib.MLP_PARAMS, dtype=numpy.float32)
        leaves = numpy.empty((self.info()['leaves'], 4), dtype=numpy.float32)
        check(lib.mpsort_gpu_model_params(self._h, raw.ctypes.data, leaves.ctypes.data))
        return raw, leaves

    def close(self):
        if self._h:
            lib.mpsort_gpu_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train(sorted_data, orderby=None, nsample=1 << 20, seed=1, linear_root=False, device_ptr=None):
    """Train a model on records sorted by data[orderby] (the output of a plain
    sort): nsample stratified positions with their exact ranks."""
    if device_ptr is None:
        a = numpy.ascontiguousarray(sorted_data)
        key = _key_desc(a.dtype, orderby)
        ptr, n, itemsize, flags = a.ctypes.data, len(a), a.dtype.itemsize, 0
    else:
        ptr, n, itemsize, key = device_ptr
        flags = _lib.DEVICE_PTRS
    if linear_root:
        flags |= _lib.MODEL_LINEAR_ROOT
    h = ctypes.c_void_p()
    check(lib.mpsort_gpu_model_train(ctypes.byref(h), ptr, n, itemsize, ctypes.byref(key), nsample, seed, flags))
    return CDFModel(h, key)


def mlp_grad(raw, x, y):
    """Loss and gradient of the root network computed by the GPU kernels."""
    raw = numpy.ascontiguousarray(raw, dtype=numpy.float32)
    x = numpy.ascontiguousarray(x, dtype=numpy.float32)
    y = numpy.ascontiguousarray(y, dtype=numpy.float32)
    g = numpy.empty(_lib.MLP_PARAMS, dtype=numpy.float32)
    loss = ctypes.c_float()
    check(lib.mpsort_gpu_mlp_grad(raw.ctypes.data, x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data,
                                  ctypes.byref(loss)))
    return loss.value, g


def sort_learned(data, model, orderby=None, out=None, nin=None, nout=None, flags=0):
    """Learned-path sort of host records over len(nin) virtual ranks."""
    data = numpy.ascontiguousarray(data)
    if out is None:
        out = numpy.empty_like(data)
    n = len(data)
    nin = [n] if nin is None else list(nin)
    nout = list(nin) if nout is None else list(nout)
    P = len(nin)
    key = _key_desc(data.dtype, orderby)
    a = (ctypes.c_int64 * P)(*nin)
    b = (ctypes.c_int64 * P)(*nout)
    check(lib.mpsort_gpu_newarray_learned(data.ctypes.data, a, P, out.ctypes.data, b, data.dtype.itemsize,
                                          ctypes.byref(key), model._h, flags))
    return out
