"""CPU checkers for the MI355X histogram sort.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only to check or time against the GPU path; nothing in the
product (mp-sort_amd/) imports it.

  restatement  liboracle.so, built from oracle.c: a C restatement of MP-sort's
               histogram sort (radix mapping, compare/bisect, stable local
               sort, piter bisection, _histogram counts, layout solve,
               exchange, second sort, checksum), each function citing the
               reference file:line it follows.
  reference    _ref/libmpsort_ref.so, _ref/ref_mpi: the reference's own C
               (radixsort.c, stdlib/msort.c, mpsort-omp.c, mpsort-mpi.c,
               mp-mpiu.c) compiled from the reference tree by oracle/Makefile.
"""
from .pyoracle import *  # noqa: F401,F403
