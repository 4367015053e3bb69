"""The C-ABI libraries load and export every function the public headers
declare (include/mpsort_gpu.h, include/mpsort.h).  CPU only: nothing here
calls into the GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "mp-sort_amd", "libmpsort_gpu.so")
MPILIB = os.path.join(ROOT, "mp-sort_amd", "libmpsort_mpi_gpu.so")

DECL = re.compile(r"^(?!typedef|#|\s|/|\*)[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def declared(header):
    """Function names declared at file scope in a header, split into the
    ones outside and inside `#ifdef MPI_VERSION`."""
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    plain, mpi, depth_mpi = [], [], False
    for line in text.splitlines(keepends=True):
        if line.startswith("#ifdef MPI_VERSION"):
            depth_mpi = True
        elif depth_mpi and line.startswith("#endif"):
            depth_mpi = False
        m = DECL.match(line)
        if m and not line.startswith("#define"):
            (mpi if depth_mpi else plain).append(m.group(1))
    return plain, mpi


def test_headers_declare_the_expected_entry_points():
    g, gm = declared("mpsort_gpu.h")
    r, rm = declared("mpsort.h")
    assert "mpsort_gpu_newarray" in g and "mpsort_gpu_newarray_learned" in g and "mpsort_gpu_model_train" in g
    assert {"radix_sort", "mpsort_omp", "mpsort_mpi_set_options", "mpsort_mpi_has_options",
            "mpsort_mpi_unset_options", "mpsort_mpi_report_last_run"} <= set(r)
    assert set(rm) == {"mpsort_mpi_impl", "mpsort_mpi_newarray_impl"}
    assert not gm


@pytest.mark.parametrize("header", ["mpsort_gpu.h", "mpsort.h"])
def test_library_exports_every_declared_function(header):
    assert os.path.exists(LIB), "build first: make -C mp-sort_amd"
    lib = ctypes.CDLL(LIB)
    plain, _ = declared(header)
    missing = [f for f in plain if not hasattr(lib, f)]
    assert not missing, missing


def test_mpi_library_exports_the_mpi_entry_points():
    if not os.path.exists(MPILIB):
        pytest.skip("no MPI compiler wrapper here: libmpsort_mpi_gpu.so is not built")
    lib = ctypes.CDLL(MPILIB)
    _, mpi = declared("mpsort.h")
    missing = [f for f in mpi if not hasattr(lib, f)]
    assert not missing, missing
