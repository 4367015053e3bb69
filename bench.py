#!/usr/bin/env python3
"""MI355X histogram sort benchmark.

Workload: 2^28 records of {u64 Morton key, i64 index} (16 B, 4 GiB) from the
clustered 3-D field (2^16 Gaussian clumps, log-normal radii sigma_ln=1,
median 2^-8, 30% uniform background, 21 bits per axis), on 64 virtual ranks
of 4,194,304 records each.  One step = one mpsort_gpu_newarray call on
device-resident input (the whole histogram sort: FirstSort, Pmin/Pmax,
splitter bisection, layout, exchange, SecondSort, checksums).

Prints one JSON line: Mrecords/s, the roofline of the dominant kernel (the
onesweep digit pass, 2*n*16 algorithmic bytes per launch), and the reference
CPU sort (mpsort_omp compiled from the reference sources, oracle/_ref) timed
on a bounded sample of the same workload on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mp-sort_amd"))
sys.path.insert(0, ROOT)

METRIC = "Mrecords/s at 2^28×16B: learned-splitter vs plain GPU sort; % of HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(recs, nsample):
    """Reference mpsort_omp (oracle/_ref, -O3) on the first nsample records."""
    import numpy
    import oracle
    if not oracle.ref_available():
        return None
    threads = min(16, os.cpu_count() or 1)
    sample = numpy.ascontiguousarray(recs[:nsample])
    key = oracle.key_for(sample.dtype, 'key')
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.ref_mpsort_omp(sample, key, nthreads=threads)
        times.append(time.perf_counter() - t0)
    times.sort()
    t = times[1]
    return {"value": round(nsample / t / 1e6, 3), "unit": "Mrecords/s", "cores": threads,
            "kind": "reference",
            "sample": "mpsort_omp (reference radixsort.c + mpsort-omp.c, gcc -O3 -fopenmp) on the "
                      "first %d records of the same clustered 16 B workload, median of 3, %.3f s"
                      % (nsample, t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--vranks", type=int, default=64)
    ap.add_argument("--cpu-sample-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain copy kernel of n*16 bytes (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the P virtual ranks of the histogram sort live "
              "on that GPU (use --vranks), and no multi-GPU figure is produced", file=sys.stderr)
        return 2

    import numpy
    from mpsort import _lib, workloads
    lib = _lib.lib
    if _lib.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 3

    n = 1 << args.log2n
    P = args.vranks
    t0 = time.perf_counter()
    recs = workloads.clustered_records(n, seed=2)
    log("generated %d clustered records in %.2f s" % (n, time.perf_counter() - t0))

    nb = recs.nbytes
    din = lib.mpsort_gpu_dmalloc(nb)
    dout = lib.mpsort_gpu_dmalloc(nb)
    if not din or not dout:
        print("bench.py: device allocation failed", file=sys.stderr)
        return 4
    _lib.check(lib.mpsort_gpu_memcpy_h2d(din, recs.ctypes.data, nb))
    cnt = [n * (t + 1) // P - n * t // P for t in range(P)]
    nin = (ctypes.c_int64 * P)(*cnt)
    key = _lib.KeyDesc(offset=0, width=8, is_signed=0, nwords=1, reserved=0)
    flags = _lib.DEVICE_PTRS | _lib.TIMING

    def step():
        _lib.check(lib.mpsort_gpu_newarray(din, nin, P, dout, nin, 16, ctypes.byref(key), flags))

    for i in range(args.warmup):
        step()
        log("warmup %d done" % i)
    _lib.check(lib.mpsort_gpu_synchronize())
    pass_ms, phases = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        st = _lib.last_stats()
        pass_ms.extend(st.pass_ms[j] for j in range(st.npass_ms))
        phases.append(st.as_dict())
    _lib.check(lib.mpsort_gpu_synchronize())
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = n / (ms * 1e-3) / 1e6

    if args.copy_probe:
        _lib.check(lib.mpsort_gpu_dcopy(dout, din, nb))
        log("copy probe: copy_kernel<4> moved %d bytes in and %d bytes out" % (nb, nb))

    # correctness of the timed output (size-independent properties)
    if args.copy_probe:
        step()
    out = numpy.empty_like(recs)
    _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, dout, nb))
    k = out['key']
    ok = bool(numpy.all(k[1:] >= k[:-1]))
    ties = k[1:] == k[:-1]
    ok = ok and bool(numpy.all(out['idx'][1:][ties] > out['idx'][:-1][ties]))
    del out
    lib.mpsort_gpu_dfree(din)
    lib.mpsort_gpu_dfree(dout)
    if not ok:
        print("bench.py: output is not the stable sort", file=sys.stderr)
        return 5

    last = phases[-1]
    alg_bytes = 2.0 * n * 16
    avg_pass_ms = sum(pass_ms) / len(pass_ms)
    achieved = alg_bytes / (avg_pass_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "onesweep_kernel<4> (one 8-bit digit pass, 2*n*16 B per launch)",
                "avg_launch_ms": round(avg_pass_ms, 4), "launches": len(pass_ms)}
    log("phases (ms): first %.3f findP %.3f lay %.3f exch %.3f second %.3f total %.3f; passes %d+%d, "
        "bisect rounds %d" % (last["t_first_sort_ms"], last["t_findp_ms"], last["t_laysolve_ms"],
                              last["t_exchange_ms"], last["t_second_sort_ms"], last["t_total_ms"],
                              last["first_passes"], last["second_passes"], last["bisect_rounds"]))
    log("pass ms: " + " ".join("%.3f" % x for x in last["pass_ms"]))
    exch_gbs = alg_bytes / (last["t_exchange_ms"] * 1e-3) / 1e9 if last["t_exchange_ms"] > 0 else 0
    log("exchange %.0f GB/s; sort-equivalent 2ns/t = %.0f GB/s" % (exch_gbs, alg_bytes / (ms * 1e-3) / 1e9))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(recs, 1 << min(args.cpu_sample_log2, args.log2n))
    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrecords/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "plain GPU histogram sort, 2^%d x 16 B {u64 Morton key, i64 index}, "
                               "clustered field, %d virtual ranks" % (args.log2n, P),
                   "records": n, "record_bytes": 16, "vranks": P, "path": "plain"},
        "phases_ms": {k2: round(last[k2], 4) for k2 in ("t_first_sort_ms", "t_findp_ms", "t_laysolve_ms",
                                                          "t_exchange_ms", "t_second_sort_ms", "t_total_ms")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
