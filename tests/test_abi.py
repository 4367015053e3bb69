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


def test_mpiu_header_declares_the_allocator_hook():
    plain, mpi = declared("mp-mpiu.h")
    assert set(plain) == {"mpiu_malloc", "mpiu_free", "mpiu_set_malloc"}
    assert set(mpi) == {"MPIU_Set_verbose_malloc"}


@pytest.mark.parametrize("header", ["mpsort_gpu.h", "mpsort.h", "mp-mpiu.h"])
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
    _, mpiu = declared("mp-mpiu.h")
    missing = [f for f in mpi + mpiu if not hasattr(lib, f)]
    assert not missing, missing


CALLER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mpsort.h"
#include "mp-mpiu.h"

/* an allocator installed the way the reference binding installs its own
 * (mpsort/binding.pyx:30-41) */
static int nalloc, nfree;
static void *my_malloc(const char *name, size_t size, const char *file, const int line, void *ud) {
    nalloc++;
    return malloc(size);
}
static void my_free(void *p, const char *file, const int line, void *ud) {
    nfree++;
    free(p);
}

int main(void) {
    MPIU_SetMalloc(my_malloc, my_free, NULL);
    void *p = MPIU_Malloc("probe", 8, 4);
    MPIU_Free(p);
    mpsort_mpi_unset_options(-1);
    mpsort_mpi_set_options(MPSORT_REQUIRE_GATHER_SORT);
    const int has = mpsort_mpi_has_options(MPSORT_REQUIRE_GATHER_SORT) != 0;
    const int hasnt = mpsort_mpi_has_options(MPSORT_DISABLE_SPARSE_ALLTOALLV) == 0;
    mpsort_mpi_unset_options(-1);
    mpiu_set_malloc(NULL, NULL, NULL);
    printf("alloc=%d free=%d options=%d\n", nalloc, nfree, has && hasnt);
    return !(nalloc == 1 && nfree == 1 && has && hasnt);
}
"""


def test_unchanged_c_caller_links_against_both_headers(tmp_path):
    """A C caller written against the reference's mpsort.h and mp-mpiu.h
    compiles, links against libmpsort_gpu.so and runs its allocator hook and
    option calls (none of which touches the GPU)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "caller.c"
    src.write_text(CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(LIB)
    subprocess.check_call([cc, "-Wall", "-Werror", "-I" + INC, "-o", str(exe), str(src), "-L" + libdir,
                           "-lmpsort_gpu", "-Wl,-rpath," + libdir])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "alloc=1 free=1 options=1" in out.stdout
