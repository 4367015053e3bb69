"""The MPI-shaped drop-in (libmpsort_mpi_gpu.so) behind an unchanged C
caller written against the reference API (mp-sort_amd/examples): the output
of mpsort_mpi / mpsort_mpi_newarray over several MPI ranks (or one rank with
virtual ranks) is the globally stable order, cut by the output counts."""
import os
import shutil
import subprocess
import sys

import numpy
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mp-sort_amd", "examples", "mpi_sort_example")
REC = numpy.dtype([('key', '<u8'), ('index', '<u8')])


def splitmix64(z):
    with numpy.errstate(over='ignore'):
        z = z + numpy.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> numpy.uint64(30))) * numpy.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> numpy.uint64(27))) * numpy.uint64(0x94D049BB133111EB)
        return z ^ (z >> numpy.uint64(31))


def expected(ranks, nbase, seed, shift):
    total = sum(nbase + 37 * r for r in range(ranks))
    g = numpy.arange(total, dtype=numpy.uint64)
    keys = splitmix64(numpy.uint64(seed) + g) >> numpy.uint64(shift)
    order = numpy.argsort(keys, kind='stable')
    out = numpy.empty(total, REC)
    out['key'] = keys[order]
    out['index'] = g[order]
    return out


@pytest.mark.parametrize("ranks,vranks,shift,inplace", [(1, 4, 0, 0), (4, 1, 0, 0), (4, 1, 44, 1), (3, 1, 58, 0)])
def test_mpi_dropin_stable_order(tmp_path, ranks, vranks, shift, inplace):
    if not os.path.exists(EXE):
        pytest.skip("MPI example not built (no MPI compiler wrapper where the tree was built)")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not os.path.exists(mpiexec):
        pytest.skip("no mpiexec")
    nbase, seed = 50000, 12345
    out = tmp_path / "out.bin"
    env = dict(os.environ, MPSORT_GPU_VRANKS=str(vranks))
    print("mpi dropin ranks=%d vranks=%d shift=%d inplace=%d" % (ranks, vranks, shift, inplace), flush=True)
    r = subprocess.run([mpiexec, "-n", str(ranks), EXE, str(nbase), str(seed), str(shift), str(inplace), str(out)],
                       env=env, timeout=120, capture_output=True, text=True)
    sys.stdout.write(r.stdout[-2000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0
    got = numpy.fromfile(out, dtype=REC)
    assert numpy.array_equal(got, expected(ranks, nbase, seed, shift))
