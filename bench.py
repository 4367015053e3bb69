#!/usr/bin/env python3
"""MI355X histogram sort benchmark.

Workload: 2^28 records of {u64 Morton key, i64 index} (16 B, 4 GiB) from the
clustered 3-D field (2^16 Gaussian clumps, log-normal radii sigma_ln=1,
median 2^-8, 30% uniform background, 21 bits per axis), on 64 virtual ranks
of 4,194,304 records each.  One step = one mpsort_gpu_newarray call on
device-resident input (the whole histogram sort: FirstSort, Pmin/Pmax,
splitter bisection, layout, exchange, SecondSort, checksums), timed for the
plain radix path and for the learned-splitter path (model trained on the GPU
beforehand, outside the timed steps); `value` is the learned path's rate.

Prints one JSON line: Mrecords/s, the roofline of the dominant kernel of the
reported path (2*n*16 algorithmic bytes per launch), and the reference CPU
sort (mpsort_omp compiled from the reference sources, oracle/_ref) timed on a
bounded sample of the same workload on this host.
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
PHASES = ("t_first_sort_ms", "t_findp_ms", "t_laysolve_ms", "t_exchange_ms", "t_second_sort_ms", "t_total_ms")
# HBM bytes per launch at 2^28 x 16 B from rocprofv3 FETCH_SIZE + WRITE_SIZE,
# each scaled by the counter's reading for a plain copy kernel of known size
# (profiles/histsort_clustered_2e28x16B_hbm_traffic.md)
TRAFFIC_ONESWEEP = 9117058048
TRAFFIC_BUCKET_SORT = 11778213312
TRAFFIC_PLAIN_BUCKET_SORT = 7984629536
# average launch time of the same kernels over the same launches (the timed
# steps of `python bench.py --steps 20 --warmup 5`) from its rocprofv3 kernel
# trace (profiles/histsort_clustered_2e28x16B_kernel_trace.md)
ROCPROF_SOURCE = "profiles/histsort_clustered_2e28x16B_kernel_trace.md"
ROCPROF_MS = {"bucket_sort_A": 2.8528, "plain_bucket_sort": 2.4325, "onesweep": 2.3646}
# per-side stages of the learned local sort, as timed by the library; with
# the merged SecondSort also its two gather-and-sort launches alone
# (merge_B: groups of buckets, then the larger single buckets), which
# bucket_sort_B contains together with the grouping of the buckets
LEARNED_KERNELS = ["hist_A", "scatter_A", "bucket_sort_A", "lds_radix_A", "fallback_A",
                   "hist_B", "scatter_B", "bucket_sort_B", "lds_radix_B", "fallback_B",
                   "merge_B"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """(threads used, logical CPUs of the host, CPUs this process may use)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    # the GPU machine gives one GPU's job a 16-thread share of its host
    threads = env if env > 0 else min(16, avail)
    return threads, os.cpu_count() or 1, avail


def cpu_baseline(recs, full=True):
    """The reference on this host's cores, on the same records.

    value: mpsort_omp (radixsort.c + mpsort-omp.c compiled with -O3 -fopenmp
    from the reference sources, oracle/_ref) on all records, OMP threads =
    the job's host share, median of 3.  Beside it: the reference Makefile's
    default build (-g -Wall -fopenmp, no optimisation), one run on all
    records; and mpirun of mpsort_mpi_newarray (mpsort-mpi.c + mp-mpiu.c)
    with one rank per thread on a 2^26-record sample, median of 3 (barrier +
    MPI_Wtime around the call, as bench-mpi.c:205-214)."""
    import numpy
    import oracle
    if not oracle.ref_available():
        return None
    threads, ncpu, avail = host_cores()
    n = len(recs)
    key = oracle.key_for(recs.dtype, 'key')
    times = []
    for _ in range(3):
        times.append(oracle.ref_mpsort_omp(recs, key, nthreads=threads, timed=True)[1])
        log("cpu baseline: mpsort_omp -O3 %d threads on %d records: %.3f s" % (threads, n, times[-1]))
    times.sort()
    t = times[1]
    out = {"value": round(n / t / 1e6, 3), "unit": "Mrecords/s", "cores": threads, "kind": "reference",
           "sample": "mpsort_omp (reference radixsort.c + mpsort-omp.c, gcc -O3 -fopenmp) on all %d records of "
                     "the same clustered 16 B workload, %d OpenMP threads (the job's host share; the host has %d "
                     "logical CPUs, %d available to this process), median of 3: %.3f s" % (n, threads, ncpu, avail, t),
           "host_logical_cpus": ncpu, "host_cpus_available": avail}
    if full and oracle.ref_available(default_flags=True):
        td = oracle.ref_mpsort_omp(recs, key, nthreads=threads, default_flags=True, timed=True)[1]
        log("cpu baseline: mpsort_omp default build on %d records: %.3f s" % (n, td))
        out["default_build"] = {"value": round(n / td / 1e6, 3), "s": round(td, 3),
                                "build": "reference Makefile default CFLAGS -g -Wall -fopenmp (no -O)",
                                "records": n, "threads": threads}
    if full and oracle.mpi_available():
        m = min(n, 1 << 26)
        sample = numpy.ascontiguousarray(recs[:m])
        cnt = [m * (t2 + 1) // threads - m * t2 // threads for t2 in range(threads)]
        try:
            sorted_m, tm = oracle.ref_mpi_sort(sample, cnt, cnt, key, repeat=3, timeout=300)
            ok = bool(numpy.all(sorted_m['key'][1:] >= sorted_m['key'][:-1]))
            log("cpu baseline: mpirun -n %d mpsort_mpi on %d records: %.3f s" % (threads, m, tm))
            out["mpi"] = {"value": round(m / tm / 1e6, 3), "s": round(tm, 4), "ranks": threads, "records": m,
                          "sorted": ok,
                          "build": "mpsort-mpi.c + mp-mpiu.c + radixsort.c, mpicc -O3 (MPICH)"}
        except Exception as e:  # MPI runtime problems are reported, not fatal
            out["mpi"] = {"error": str(e)[-300:]}
    return out


def check_stable_sorted(out, n):
    """sorted by key, ties in input (idx) order, and a permutation of 0..n-1"""
    import numpy
    k = out['key']
    ok = bool(numpy.all(k[1:] >= k[:-1]))
    ties = k[1:] == k[:-1]
    ok = ok and bool(numpy.all(out['idx'][1:][ties] > out['idx'][:-1][ties]))
    del ties
    seen = numpy.zeros(n, dtype=bool)
    seen[out['idx']] = True
    return ok and bool(seen.all())


def bench_batched(args):
    """The batched workload: 10,000 independent stable sorts, sizes
    log-uniform on [1e3, 1e7] rescaled to 2^30 records in total, each segment
    of one of two record types chosen by a seeded draw ({i32 key, f32 value}
    or {i64 key, i64 value}), keys uniform over the full signed range.  One
    plan (both types) captured in a hipGraph, tuned on its first run, against
    the per-segment launch loop (every segment by launches of its own, no
    host wait between them) and a loop of one sort call per segment."""
    import numpy
    from mpsort import _lib, segments, workloads
    lib = _lib.lib
    nseg = args.segments
    total = 1 << args.log2n
    sizes = workloads.segment_sizes(nseg, 4, lo=1e3, hi=1e7, total=total).astype(numpy.int64)
    pick = numpy.array([workloads.mix64(0xB0A7 + i) & 1 for i in range(nseg)], dtype=bool)
    groups = []
    for gi, (dt, sel) in enumerate(((workloads.REC8, ~pick), (workloads.REC16I, pick))):
        sz = sizes[sel]
        n = int(sz.sum())
        a = numpy.zeros(n, dt)
        raw = workloads.uniform(n, 40 + gi)
        if dt == workloads.REC8:
            a['key'] = (raw >> numpy.uint64(32)).astype(numpy.uint32).view(numpy.int32)
            # the f32 value field carries the record's position as its bit pattern
            a['value'] = numpy.arange(n, dtype=numpy.uint32).view(numpy.float32)
        else:
            a['key'] = raw.view(numpy.int64)
            a['value'] = numpy.arange(n, dtype=numpy.int64)
        del raw
        src = lib.mpsort_gpu_dmalloc(a.nbytes)
        dst = lib.mpsort_gpu_dmalloc(a.nbytes)
        _lib.check(lib.mpsort_gpu_memcpy_h2d(src, a.ctypes.data, a.nbytes))
        groups.append(dict(dtype=dt, sizes=sz, n=n, host=a, src=src, dst=dst))
        log("group %d: %d segments, %d records of %d B" % (gi, len(sz), n, dt.itemsize))
    nrec = sum(g["n"] for g in groups)
    nbytes = sum(g["n"] * g["dtype"].itemsize for g in groups)

    plan = segments.SegmentPlan(graph=True, tune=True)
    for g in groups:
        plan.add(g["src"], g["dst"], g["sizes"], g["dtype"], 'key')
    t0 = time.perf_counter()
    plan.run()  # tunes, captures
    plan_info = [plan.info(i) for i in range(2)]
    log("plan tuned and captured in %.2f s: %s" % (time.perf_counter() - t0, plan_info))
    for i in range(args.warmup):
        plan.run()
    _lib.check(lib.mpsort_gpu_synchronize())
    t0 = time.perf_counter()
    dev_ms = []
    for i in range(args.steps):
        dev_ms.append(plan.run())
    _lib.check(lib.mpsort_gpu_synchronize())
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log("batched: %.3f ms per run (device %.3f ms)" % (ms, sum(dev_ms) / len(dev_ms)))

    # correctness: sorted, stable (values are the input positions) and a
    # permutation of each segment
    for g in groups:
        out = numpy.empty_like(g["host"])
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, g["dst"], out.nbytes))
        starts = numpy.concatenate([[0], numpy.cumsum(g["sizes"])[:-1]])
        seg_of = numpy.repeat(numpy.arange(len(g["sizes"])), g["sizes"])
        k = out['key']
        v = (out['value'].view(numpy.uint32) if g["dtype"] == workloads.REC8 else out['value']).astype(numpy.int64)
        same = seg_of[1:] == seg_of[:-1]
        ok = bool(numpy.all(~same | (k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (v[1:] > v[:-1]))))
        ok = ok and bool(numpy.all((v >= starts[seg_of]) & (v < (starts + g["sizes"])[seg_of])))
        seen = numpy.zeros(g["n"], dtype=bool)
        seen[v] = True
        ok = ok and bool(seen.all())
        del out, seg_of, seen
        if not ok:
            print("bench.py: batched output is not the per-segment stable sort", file=sys.stderr)
            return 5
    plan.close()

    if args.plan_only:
        for g in groups:
            lib.mpsort_gpu_dfree(g["src"])
            lib.mpsort_gpu_dfree(g["dst"])
        print(json.dumps({"metric": "Mrecords/s, batched segmented sort (%d segments, plan only)" % nseg,
                          "value": round(nrec / (ms * 1e-3) / 1e6, 2), "unit": "Mrecords/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i32/i64",
                          "data": "synthetic", "config": {"workload": "batched, plan only", "segments": nseg,
                                                          "records": nrec, "bytes": nbytes},
                          "tuned_plan": {"i32_f32": plan_info[0], "i64_i64": plan_info[1]}}), flush=True)
        return 0

    # the per-segment loop the plan replaces, with the same kernels, tile
    # shapes and workspaces: every segment by launches of its own, enqueued
    # back to back (no host wait between segments; prepared outside the
    # timing, like the plan)
    each = segments.SegmentPlan(each=True)
    for g in groups:
        each.add(g["src"], g["dst"], g["sizes"], g["dtype"], 'key')
    each.run()
    _lib.check(lib.mpsort_gpu_synchronize())
    loop_steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for i in range(loop_steps):
        each.run()
    _lib.check(lib.mpsort_gpu_synchronize())
    each_s = (time.perf_counter() - t0) / loop_steps
    each.close()
    log("per-segment launch loop: %.3f s per run" % each_s)

    # one plain sort call per segment, each waiting for its result
    t0 = time.perf_counter()
    for g in groups:
        key = _lib.KeyDesc(offset=0, width=g["dtype"]['key'].itemsize, is_signed=1, nwords=1, reserved=0)
        es = g["dtype"].itemsize
        off = 0
        for c in g["sizes"]:
            c = int(c)
            cnt = (ctypes.c_int64 * 1)(c)
            _lib.check(lib.mpsort_gpu_newarray(g["src"] + off * es, cnt, 1, g["dst"] + off * es, cnt, es,
                                               ctypes.byref(key), _lib.DEVICE_PTRS | _lib.NO_CHECKSUM))
            off += c
    _lib.check(lib.mpsort_gpu_synchronize())
    loop_s = time.perf_counter() - t0
    log("per-call loop: %.3f s" % loop_s)
    for g in groups:
        lib.mpsort_gpu_dfree(g["src"])
        lib.mpsort_gpu_dfree(g["dst"])

    value = nrec / (ms * 1e-3) / 1e6
    loop_value = nrec / loop_s / 1e6
    each_value = nrec / each_s / 1e6
    # bytes the plan's passes move: a segment up to the in-LDS threshold is
    # read and written once; a larger one is read by the digit histogram and
    # read and written by one onesweep pass per key byte (uniform keys leave
    # no digit place trivial)
    pass_bytes = 0
    for g, info in zip(groups, plan_info):
        es = g["dtype"].itemsize
        kb = g["dtype"]['key'].itemsize
        small = int(g["sizes"][g["sizes"] <= info["threshold"]].sum())
        pass_bytes += small * 2 * es + (g["n"] - small) * es * (1 + 2 * kb)
    pass_gbs = pass_bytes / (ms * 1e-3) / 1e9
    result = {
        "metric": "Mrecords/s, batched segmented sort (10,000 segments, 2^30 records)", "value": round(value, 2),
        "unit": "Mrecords/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i32/i64", "data": "synthetic",
        "config": {"workload": "batched: %d segments log-uniform [1e3, 1e7] rescaled to %d records, "
                               "{i32,f32} and {i64,i64} records" % (nseg, nrec),
                   "segments": nseg, "records": nrec, "bytes": nbytes},
        "per_segment_loop": {"value": round(each_value, 2), "s": round(each_s, 3), "runs": loop_steps,
                             "form": "every segment by launches of its own on the stream, back to back, no host "
                                     "wait between segments (MPSORT_GPU_SEGPLAN_EACH, untuned default shapes)"},
        "speedup_vs_loop": round(value / each_value, 2),
        "per_call_loop": {"value": round(loop_value, 2), "s": round(loop_s, 3),
                          "form": "one mpsort_gpu_newarray call per segment, each waiting for its result"},
        "sort_equivalent_GBps": round(2.0 * nbytes / (ms * 1e-3) / 1e9, 1),
        "passes": {"bytes": pass_bytes, "GBps": round(pass_gbs, 1), "frac_of_hbm_peak": round(pass_gbs / HBM_PEAK_GBS, 3),
                   "form": "segments up to the in-LDS threshold: one read and one write; larger ones: the "
                           "histogram read plus one read and one write per key byte (4 passes for i32 keys, "
                           "8 for i64)"},
        "tuned_plan": {"i32_f32": plan_info[0], "i64_i64": plan_info[1]},
    }
    print(json.dumps(result), flush=True)
    return 0


def bench_snapshot(args):
    """The snapshot-matching workload: two arrays of 2^27 particles
    ({u64 id, f32 pos[3], f32 vel[3]}, 32 B): A with the ids a seeded
    permutation, B = A sorted by id with 0.1% of the records displaced (its
    exchange is sparse).  Both sorted by id on 64 virtual ranks, learned path
    with one model (trained on A's plain sort) for both; one step = the sorts
    of A and B.  Beside it, the plain path's exchange with the dense and the
    sparse run table, for A and for B, on 64 and on 4096 virtual ranks."""
    import numpy
    from mpsort import _lib, learned, workloads
    lib = _lib.lib
    n, P = 1 << args.log2n, args.vranks
    t0 = time.perf_counter()
    a, b = workloads.snapshot_pair(n, 27)
    log("generated 2 x %d particles in %.2f s" % (n, time.perf_counter() - t0))
    key = learned._key_desc(a.dtype, 'id')

    def counts(p):
        return (ctypes.c_int64 * p)(*[n * (t + 1) // p - n * t // p for t in range(p)])

    nin = counts(P)
    da, db, dout = (lib.mpsort_gpu_dmalloc(a.nbytes) for _ in range(3))
    if not (da and db and dout):
        print("bench.py: device allocation failed", file=sys.stderr)
        return 4
    _lib.check(lib.mpsort_gpu_memcpy_h2d(da, a.ctypes.data, a.nbytes))
    _lib.check(lib.mpsort_gpu_memcpy_h2d(db, b.ctypes.data, b.nbytes))
    del b
    flags = _lib.DEVICE_PTRS | _lib.TIMING

    def timed(step, label, steps=None):
        steps = steps or args.steps
        for i in range(args.warmup):
            step()
        _lib.check(lib.mpsort_gpu_synchronize())
        t0 = time.perf_counter()
        for i in range(steps):
            step()
        _lib.check(lib.mpsort_gpu_synchronize())
        ms = (time.perf_counter() - t0) * 1e3 / steps
        log("%s: %.3f ms per step" % (label, ms))
        return ms

    # the plain path's exchange: dense and sparse run tables, A and B
    exchange = {}
    for p in (P, 4096):
        cp = counts(p)
        for name, src in (("A", da), ("B", db)):
            for pol, f in (("dense", _lib.EXCHANGE_DENSE), ("sparse", _lib.EXCHANGE_SPARSE)):
                def one():
                    _lib.check(lib.mpsort_gpu_newarray(src, cp, p, dout, cp, 32, ctypes.byref(key), flags | f))
                ms = timed(one, "plain %s P=%d %s exchange" % (name, p, pol), steps=max(2, args.steps // 2))
                st = _lib.last_stats()
                exchange.setdefault("P%d" % p, {}).setdefault(name, {})[pol] = {
                    "sort_ms": round(ms, 3), "exchange_ms": round(st.t_exchange_ms, 4),
                    "runs_walked": st.exchange_runs, "sparse": st.exchange_sparse}

    def plain_step():
        for src in (da, db):
            _lib.check(lib.mpsort_gpu_newarray(src, nin, P, dout, nin, 32, ctypes.byref(key), flags))

    ms_plain = timed(plain_step, "plain (A then B)")
    ref = numpy.empty_like(a)
    _lib.check(lib.mpsort_gpu_newarray(da, nin, P, dout, nin, 32, ctypes.byref(key), flags))
    _lib.check(lib.mpsort_gpu_memcpy_d2h(ref.ctypes.data, dout, ref.nbytes))
    if not numpy.array_equal(ref['id'], numpy.arange(n, dtype=numpy.uint64)):
        print("bench.py: plain output ids are not 0..n-1", file=sys.stderr)
        return 5
    t0 = time.perf_counter()
    model = learned.train(None, nsample=1 << 20, seed=1, device_ptr=(dout, n, 32, key))
    t_train = time.perf_counter() - t0

    def learned_step():
        for src in (da, db):
            _lib.check(lib.mpsort_gpu_newarray_learned(src, nin, P, dout, nin, 32, ctypes.byref(key), model._h,
                                                       flags))

    ms = timed(learned_step, "learned (A then B)")
    st = _lib.last_stats().as_dict()
    ok = True
    out = numpy.empty_like(a)
    for src in (da, db):
        _lib.check(lib.mpsort_gpu_newarray_learned(src, nin, P, dout, nin, 32, ctypes.byref(key), model._h, flags))
        _lib.check(lib.mpsort_gpu_memcpy_d2h(out.ctypes.data, dout, out.nbytes))
        # B holds A's records (0.1% of them moved): both sort to the same array
        ok = ok and bool(numpy.array_equal(out.view(numpy.uint8), ref.view(numpy.uint8)))
    model.close()
    for p in (da, db, dout):
        lib.mpsort_gpu_dfree(p)
    if not ok:
        print("bench.py: learned snapshot outputs differ from the plain sort of A", file=sys.stderr)
        return 6
    value = 2 * n / (ms * 1e-3) / 1e6
    result = {
        "metric": "Mrecords/s, snapshot matching (2 x 2^%d x 32 B particles sorted by id, learned splitters)"
                  % args.log2n,
        "value": round(value, 2), "unit": "Mrecords/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "snapshot matching: A = seeded permutation of ids with uniform pos/vel, "
                               "B = A sorted by id with 0.1%% of records displaced, %d virtual ranks" % P,
                   "records_per_array": n, "record_bytes": 32},
        "ids_equal_and_sorted": ok, "train_s": round(t_train, 3),
        "plain": {"value": round(2 * n / (ms_plain * 1e-3) / 1e6, 2), "ms_per_step": round(ms_plain, 3)},
        "learned_vs_plain": round(ms_plain / ms, 3),
        "last_sort_phases_ms": {k: round(st[k], 4) for k in PHASES},
        "sort_equivalent_GBps": round(2 * 2.0 * n * 32 / (ms * 1e-3) / 1e9, 1),
        "plain_exchange": exchange,
    }
    print(json.dumps(result), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--vranks", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="time the plain path only")
    ap.add_argument("--host-leg", action="store_true",
                    help="also time one call on host buffers (PCIe transfers included)")
    ap.add_argument("--workload", default="histsort", choices=["histsort", "batched", "snapshot", "dists"],
                    help="histsort: the 2^28 x 16 B histogram sort (default); batched: 10,000 independent "
                         "sorts; snapshot: two 2^27 x 32 B particle snapshots sorted by id; dists: the "
                         "learned path on config 3's four key distributions")
    ap.add_argument("--dists", default=",".join(DISTS))
    ap.add_argument("--cpu-quick", action="store_true",
                    help="CPU baseline: the -O3 mpsort_omp only (no default build, no mpirun)")
    ap.add_argument("--segments", type=int, default=10000)
    ap.add_argument("--plan-only", action="store_true",
                    help="batched workload: time the plan only (no per-segment or per-call loop), "
                         "e.g. to keep a kernel trace of it small")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain copy kernel of n*16 bytes (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the P virtual ranks of the histogram sort live "
              "on that GPU (use --vranks), and no multi-GPU figure is produced", file=sys.stderr)
        return 2

    from mpsort import _lib
    if _lib.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 3
    if args.workload == "batched":
        if args.log2n == 28:
            args.log2n = 30
        return bench_batched(args)
    if args.workload == "snapshot":
        if args.log2n == 28:
            args.log2n = 27
        return bench_snapshot(args)
    if args.workload == "dists":
        return bench_dists(args)

    return bench_histsort(args)


def _device_copy(recs):
    from mpsort import _lib
    lib = _lib.lib
    d = lib.mpsort_gpu_dmalloc(recs.nbytes)
    if not d:
        raise MemoryError("device allocation of %d bytes failed" % recs.nbytes)
    _lib.check(lib.mpsort_gpu_memcpy_h2d(d, recs.ctypes.data, recs.nbytes))
    return d


def plain_first_form(st):
    """(FirstSort form, entries of pass_ms before the SecondSort's) of a
    plain-path call's statistics"""
    nsec = 4 if st["second_merged"] else st["second_passes"]
    if st["first_passes"] == 0 and len(st["pass_ms"]) - nsec == 6:
        return "buckets", 6
    return "radix", st["first_passes"]


def plain_stages(st):
    """the plain path's stage times (ms) from one call's pass_ms"""
    form, k0 = plain_first_form(st)
    ms = st["pass_ms"]
    out = {}
    if form == "buckets":
        for i, name in enumerate(("sample_tables", "hist", "scatter", "bucket_sort", "lds_radix", "fallback")):
            out["first_" + name] = round(ms[i], 4)
    else:
        out["first_radix_passes"] = round(sum(ms[:k0]), 4)
    rest = ms[k0:]
    if st["second_merged"]:
        for i, name in enumerate(("tables_bounds_groups", "gather_sort", "listed", "fallback")):
            out["second_" + name] = round(rest[i], 4)
    elif rest:
        out["second_radix_passes"] = round(sum(rest), 4)
    return out


def _timed(step, label, steps, warmup):
    """warmup + timed calls of step; (ms per step, per-call stats of the timed
    calls, per-call stats of the warmup calls)"""
    from mpsort import _lib
    lib = _lib.lib
    warm = []
    for i in range(warmup):
        step()
        warm.append(_lib.last_stats().as_dict())
        log("%s warmup %d done" % (label, i))
    _lib.check(lib.mpsort_gpu_synchronize())
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        stats.append(_lib.last_stats().as_dict())
    _lib.check(lib.mpsort_gpu_synchronize())
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, stats, warm


def train_model(n, P, key, gen, seed, label):
    """A model trained by the GPU kernels on one plain GPU sort of a training
    realisation gen(seed) (2^20 stratified samples with their exact ranks);
    returns (model, plain sort ms, training s)."""
    from mpsort import _lib, learned
    lib = _lib.lib
    t0 = time.perf_counter()
    recs = gen(n, seed)
    log("%s: training realisation (seed %d) generated in %.2f s" % (label, seed, time.perf_counter() - t0))
    din = _device_copy(recs)
    del recs
    dout = lib.mpsort_gpu_dmalloc(n * 16)
    cnt = [n * (t + 1) // P - n * t // P for t in range(P)]
    nin = (ctypes.c_int64 * P)(*cnt)
    _lib.check(lib.mpsort_gpu_newarray(din, nin, P, dout, nin, 16, ctypes.byref(key), _lib.DEVICE_PTRS))
    _lib.check(lib.mpsort_gpu_synchronize())
    t_plain = _lib.last_stats().t_total_ms
    lib.mpsort_gpu_dfree(din)
    t0 = time.perf_counter()
    model = learned.train(None, nsample=1 << 20, seed=1, device_ptr=(dout, n, 16, key))
    t_train = time.perf_counter() - t0
    lib.mpsort_gpu_dfree(dout)
    return model, t_plain, t_train


def bench_histsort(args):
    """The headline: 2^28 clustered 16 B records on 64 virtual ranks.

    The model is trained on a plain sort of ANOTHER realisation of the same
    clustered field (same clumps, independent points: field seed 2, point
    seed 3); the timed array is the realisation of point seed 2.  Reported
    beside the reuse rate: the one-shot cost of an unseen field (plain sort
    of a training realisation + training + the learned sort)."""
    import numpy
    from mpsort import _lib, workloads
    lib = _lib.lib
    n = 1 << args.log2n
    P = args.vranks
    cnt = [n * (t + 1) // P - n * t // P for t in range(P)]
    nin = (ctypes.c_int64 * P)(*cnt)
    key = _lib.KeyDesc(offset=0, width=8, is_signed=0, nwords=1, reserved=0)
    flags = _lib.DEVICE_PTRS | _lib.TIMING
    field = 2

    model = None
    if not args.no_learned:
        model, t_plain_train, t_train = train_model(
            n, P, key, lambda m, sd: workloads.clustered_records(m, seed=sd, field_seed=field), 3, "clustered")
        minfo = model.info()
        log("model: trained in %.3f s (%d Adam steps, %d leaves, err_max %.2e, max leaf %d samples)"
            % (t_train, minfo['steps'], minfo['leaves'], minfo['err_max'], minfo['max_leaf_samples']))

    t0 = time.perf_counter()
    recs = workloads.clustered_records(n, seed=2, field_seed=field)
    log("generated %d clustered records (timed realisation) in %.2f s" % (n, time.perf_counter() - t0))
    nb = recs.nbytes
    din = _device_copy(recs)
    dout = lib.mpsort_gpu_dmalloc(nb)
    if not dout:
        print("bench.py: device allocation failed", file=sys.stderr)
        return 4

    def plain_step():
        _lib.check(lib.mpsort_gpu_newarray(din, nin, P, dout, nin, 16, ctypes.byref(key), flags))

    # ---- plain path ------------------------------------------------------
    ms_plain, st_plain, _ = _timed(plain_step, "plain", args.steps, args.warmup)
    log("plain: %.3f ms per step, %.1f Mrecords/s" % (ms_plain, n / (ms_plain * 1e-3) / 1e6))

    # the reference's shape on the GPU: radix FirstSort, exchange, radix
    # SecondSort (fewer steps: it is a reported figure, not `value`)
    rflags = flags | _lib.FIRST_RADIX | _lib.EXCHANGE

    def radix_step():
        _lib.check(lib.mpsort_gpu_newarray(din, nin, P, dout, nin, 16, ctypes.byref(key), rflags))

    rsteps = max(3, args.steps // 4)
    ms_radix, st_radix, _ = _timed(radix_step, "plain radix", rsteps, min(args.warmup, 2))
    log("plain, reference shape (radix passes + exchange): %.3f ms per step" % ms_radix)
    if args.copy_probe:
        _lib.check(lib.mpsort_gpu_dcopy(dout, din, nb))
        log("copy probe: copy_kernel<4> moved %d bytes in and %d bytes out" % (nb, nb))
        plain_step()
    plain_out = numpy.empty_like(recs)
    _lib.check(lib.mpsort_gpu_memcpy_d2h(plain_out.ctypes.data, dout, nb))
    if not check_stable_sorted(plain_out, n):
        print("bench.py: plain output is not the stable sort", file=sys.stderr)
        return 5

    # ---- learned path (model from the other realisation) ----------------
    result_learned = None
    if model is not None:
        def learned_step():
            _lib.check(lib.mpsort_gpu_newarray_learned(din, nin, P, dout, nin, 16, ctypes.byref(key),
                                                       model._h, flags))

        ms_learned, st_learned, warm_learned = _timed(learned_step, "learned", args.steps, args.warmup)
        log("learned: %.3f ms per step, %.1f Mrecords/s" % (ms_learned, n / (ms_learned * 1e-3) / 1e6))
        lout = numpy.empty_like(recs)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(lout.ctypes.data, dout, nb))
        same = bool(numpy.array_equal(lout, plain_out))
        del lout
        if not same:
            print("bench.py: learned output differs from the plain path", file=sys.stderr)
            return 6
        L = st_learned[-1]
        one_shot_ms = t_plain_train + t_train * 1e3 + ms_learned
        gain = ms_plain - ms_learned
        result_learned = {
            "ms_per_step": round(ms_learned, 3),
            "value": round(n / (ms_learned * 1e-3) / 1e6, 2),
            "bit_identical_to_plain": same,
            "model_trained_on": "plain GPU sort of another realisation of the clustered field (same clumps, "
                                "independent points: field seed %d, point seed 3; timed: point seed 2)" % field,
            "train_s": round(t_train, 3), "model": {k2: minfo[k2] for k2 in minfo},
            "one_shot": {"ms": round(one_shot_ms, 2), "plain_sort_of_training_array_ms": round(t_plain_train, 3),
                         "train_ms": round(t_train * 1e3, 1), "learned_sort_ms": round(ms_learned, 3),
                         "vs_plain": round(ms_plain / one_shot_ms, 3),
                         "break_even_sorts": (int(numpy.ceil((t_plain_train + t_train * 1e3) / gain))
                                              if gain > 0 else None)},
            "splitters_exact": L["splitters_exact"],
            "splitters_repaired": L["splitters_repaired"],
            "repair_rounds": L["repair_rounds"],
            "buckets_overflow": L["buckets_overflow"],
            "buckets_overflow_first_sort": L["buckets_overflow_first"],
            "plain_sides": L["plain_sides"],
            "buckets_lds_radix": L["buckets_lds_radix"],
            "buckets_listed": L["buckets_listed"],
            "phases_ms": {k2: round(L[k2], 4) for k2 in PHASES},
            "kernels_ms": dict(zip(LEARNED_KERNELS, [round(x, 4) for x in L["pass_ms"]]))}
        log("learned phases: " + json.dumps(result_learned["phases_ms"]))
        log("learned kernels: " + json.dumps(result_learned["kernels_ms"]))
        model.close()
    lib.mpsort_gpu_dfree(din)
    lib.mpsort_gpu_dfree(dout)
    # one call with host buffers (PCIe copies in and out included): reported,
    # never the headline value
    host_leg = None
    if args.host_leg:
        t0 = time.perf_counter()
        _lib.check(lib.mpsort_gpu_newarray(recs.ctypes.data, nin, P, plain_out.ctypes.data, nin, 16,
                                           ctypes.byref(key), _lib.TIMING))
        t_host = time.perf_counter() - t0
        hs = _lib.last_stats().as_dict()
        host_leg = {"ms": round(t_host * 1e3, 3), "value": round(n / t_host / 1e6, 2), "unit": "Mrecords/s",
                    "h2d_ms": round(hs["t_h2d_ms"], 3), "d2h_ms": round(hs["t_d2h_ms"], 3),
                    "path": "plain, host pointers (pageable numpy buffers)"}
        log("host-pointer call: %.3f ms (h2d %.3f, d2h %.3f)" % (t_host * 1e3, hs["t_h2d_ms"], hs["t_d2h_ms"]))
    del plain_out

    last = st_plain[-1]
    alg_bytes = 2.0 * n * 16  # one read and one write of every record

    def roof(launch_ms, kernel, traffic, rocprof_ms, launch_bytes=alg_bytes):
        avg = sum(launch_ms) / len(launch_ms)
        achieved = launch_bytes / (avg * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
             "avg_launch_ms": round(avg, 4), "launches": len(launch_ms),
             "timing": "HIP events around each launch on the library's stream, the timed steps only "
                       "(warmup launches, the cold first one included, are left out)"}
        if rocprof_ms:
            # the committed rocprofv3 trace of the driver's command, averaged
            # over the same launches (timed steps only, warmups left out)
            r["rocprof"] = {"avg_launch_ms": rocprof_ms, "frac": round(launch_bytes / (rocprof_ms * 1e-3) / 1e9 /
                                                                        HBM_PEAK_GBS, 4),
                            "launches": "the timed steps' launches of `rocprofv3 --kernel-trace -- python3 "
                                        "bench.py --gpus 1 --steps 20 --warmup 5 --no-cpu`",
                            "source": ROCPROF_SOURCE}
        return r

    radix_roof = roof([x for st in st_radix for x in st["pass_ms"][:st["first_passes"] + st["second_passes"]]],
                      "onesweep_kernel<4,1024,8> (one 8-bit digit pass, 2*n*16 B per launch)", TRAFFIC_ONESWEEP,
                      ROCPROF_MS.get("onesweep"))
    first_form, k0 = plain_first_form(st_plain[-1])
    if first_form == "radix":
        plain_roof = roof([x for st in st_plain for x in st["pass_ms"][:st["first_passes"] + st["second_passes"]]],
                          "onesweep_kernel<4,1024,8> (one 8-bit digit pass, 2*n*16 B per launch)", TRAFFIC_ONESWEEP,
                          ROCPROF_MS.get("onesweep"))
    else:
        plain_roof = roof([st["pass_ms"][3] for st in st_plain],
                          "interp_sort_kernel<4,512,9> (the plain FirstSort's in-LDS bucket sort, sampled tables, "
                          "one launch per step over all n records: 2*n*16 B per launch)", TRAFFIC_PLAIN_BUCKET_SORT,
                          ROCPROF_MS.get("plain_bucket_sort"))
    if result_learned is not None:
        ib = LEARNED_KERNELS.index
        # the step's largest kernel: the FirstSort's in-LDS bucket sort, one
        # launch per step over all n records (the merged SecondSort splits
        # its records over two smaller kernels, merge_B in kernels_ms)
        roofline = roof([st["pass_ms"][ib("bucket_sort_A")] for st in st_learned],
                        "interp_sort_kernel<4,512,9> (FirstSort in-LDS bucket sort, one launch per step over all n "
                        "records: 2*n*16 B per launch)", TRAFFIC_BUCKET_SORT, ROCPROF_MS.get("bucket_sort_A"))
        result_learned["roofline_by_stage"] = {
            k: {"avg_ms": round(sum(st["pass_ms"][ib(k)] for st in st_learned) / len(st_learned), 4),
                "GB/s": round((alg_bytes / 2 if k.startswith("hist") else alg_bytes) /
                              (sum(st["pass_ms"][ib(k)] for st in st_learned) / len(st_learned) * 1e-3) / 1e9, 1)}
            for k in ("hist_A", "scatter_A", "bucket_sort_A", "merge_B")}
    else:
        roofline = plain_roof
    log("plain phases (ms): first %.3f findP %.3f lay %.3f exch %.3f second %.3f total %.3f; passes %d+%d, "
        "bisect rounds %d" % (last["t_first_sort_ms"], last["t_findp_ms"], last["t_laysolve_ms"],
                              last["t_exchange_ms"], last["t_second_sort_ms"], last["t_total_ms"],
                              last["first_passes"], last["second_passes"], last["bisect_rounds"]))
    log("plain pass ms: " + " ".join("%.3f" % x for x in last["pass_ms"]))
    if first_form == "buckets":
        plain_first = {
            "form": "bucket sort (tables from a sorted sample of 65536 keys per rank)",
            "stages_ms": {name: round(sum(st["pass_ms"][i] for st in st_plain) / len(st_plain), 4)
                          for i, name in enumerate(("sample_tables", "hist", "scatter", "bucket_sort", "lds_radix",
                                                    "fallback"))},
            "buckets_overflow": last["buckets_overflow_first"]}
        log("plain bucket FirstSort: %s" % json.dumps(plain_first))
    else:
        plain_first = {"form": "%d digit passes" % last["first_passes"]}
    plain_second = None
    if last["second_merged"]:
        # the SecondSort without an exchange: its stages after the FirstSort's
        plain_second = {
            "form": "merged (receiver buckets gathered from the senders' sorted segments, sampled tables)",
            "stages_ms": {name: round(sum(st["pass_ms"][k0 + i] for st in st_plain) / len(st_plain), 4)
                          for i, name in enumerate(("tables_bounds_groups", "gather_sort", "listed", "fallback"))},
            "buckets_overflow": last["buckets_overflow"] - (last["buckets_overflow_first"]
                                                            if first_form == "buckets" else 0),
            "buckets_listed_both_sorts": last["buckets_listed"],
            "buckets_lds_radix_both_sorts": last["buckets_lds_radix"], "plain_sides": last["plain_sides"]}
        log("plain merged SecondSort: %s" % json.dumps(plain_second))
    else:
        plain_second = {"form": "exchange, then %d digit passes" % last["second_passes"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(recs, full=not args.cpu_quick)
    value_plain = n / (ms_plain * 1e-3) / 1e6
    if result_learned is not None:
        value, ms, path = result_learned["value"], result_learned["ms_per_step"], "learned"
    else:
        value, ms, path = round(value_plain, 2), ms_plain, "plain"
    result = {
        "metric": METRIC, "value": value, "unit": "Mrecords/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "GPU histogram sort, 2^%d x 16 B {u64 Morton key, i64 index}, "
                               "clustered field, %d virtual ranks" % (args.log2n, P),
                   "records": n, "record_bytes": 16, "vranks": P, "path": path},
        "plain": {"value": round(value_plain, 2), "ms_per_step": round(ms_plain, 3),
                  "phases_ms": {k2: round(last[k2], 4) for k2 in PHASES}, "first_sort": plain_first,
                  "second_sort": plain_second},
        "learned": result_learned,
        "host_pointer_call": host_leg,
        "learned_vs_plain": round(value / value_plain, 3) if result_learned else None,
        "sort_equivalent_GBps": round(alg_bytes / (ms * 1e-3) / 1e9, 1),
        "roofline": roofline,
        "plain_roofline": plain_roof,
        "plain_radix": {"ms_per_step": round(ms_radix, 3), "steps": rsteps,
                        "value": round(n / (ms_radix * 1e-3) / 1e6, 2),
                        "form": "radix FirstSort (%d digit passes), exchange, radix SecondSort (%d passes)" % (
                            st_radix[-1]["first_passes"], st_radix[-1]["second_passes"]),
                        "phases_ms": {k2: round(st_radix[-1][k2], 4) for k2 in PHASES},
                        "roofline": radix_roof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    return 0


DISTS = ("uniform", "lognormal", "zipf", "nearly_sorted")


def bench_dists(args):
    """Config 3's four key distributions at the headline shape (2^28 x 16 B,
    64 virtual ranks): uniform u64, a log-space mixture of 8 Gaussians, Zipf
    s=1.1 over 2^20 values, a sorted ramp with 1% of the entries swapped in
    random pairs.  For each: the plain sort, a model trained on a plain sort
    of an independent 2^24-record realisation of the distribution, the
    learned sort (checked byte for byte against the plain output), and the
    learned path's splitter repair and fallback counts."""
    import numpy
    from mpsort import _lib, workloads
    lib = _lib.lib
    n = 1 << args.log2n
    P = args.vranks
    cnt = [n * (t + 1) // P - n * t // P for t in range(P)]
    nin = (ctypes.c_int64 * P)(*cnt)
    key = _lib.KeyDesc(offset=0, width=8, is_signed=0, nwords=1, reserved=0)
    flags = _lib.DEVICE_PTRS | _lib.TIMING
    gens = {"uniform": workloads.uniform, "lognormal": workloads.lognormal_mix, "zipf": workloads.zipf,
            "nearly_sorted": workloads.nearly_sorted}
    out = {}
    for dist in args.dists.split(","):
        g = gens[dist]
        model, _, t_train = train_model(1 << 24, P, key, lambda m, sd: workloads.records16(g(m, sd)), 101, dist)
        t0 = time.perf_counter()
        recs = workloads.records16(g(n, 5))
        log("%s: %d records generated in %.2f s" % (dist, n, time.perf_counter() - t0))
        din = _device_copy(recs)
        dout = lib.mpsort_gpu_dmalloc(recs.nbytes)

        def plain_step():
            _lib.check(lib.mpsort_gpu_newarray(din, nin, P, dout, nin, 16, ctypes.byref(key), flags))

        def learned_step():
            _lib.check(lib.mpsort_gpu_newarray_learned(din, nin, P, dout, nin, 16, ctypes.byref(key),
                                                       model._h, flags))

        ms_p, st_p, _ = _timed(plain_step, dist + " plain", args.steps, args.warmup)
        ref = numpy.empty_like(recs)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(ref.ctypes.data, dout, recs.nbytes))
        ok = check_stable_sorted(ref, n)
        ms_l, st_l, _ = _timed(learned_step, dist + " learned", args.steps, args.warmup)
        got = numpy.empty_like(recs)
        _lib.check(lib.mpsort_gpu_memcpy_d2h(got.ctypes.data, dout, recs.nbytes))
        same = bool(numpy.array_equal(got, ref))
        del got, ref, recs
        lib.mpsort_gpu_dfree(din)
        lib.mpsort_gpu_dfree(dout)
        L = st_l[-1]
        out[dist] = {"plain_ms": round(ms_p, 3), "learned_ms": round(ms_l, 3),
                     "learned_vs_plain": round(ms_p / ms_l, 3),
                     "learned_Mrecords_s": round(n / (ms_l * 1e-3) / 1e6, 1),
                     "plain_stable_sorted": ok, "learned_bit_identical_to_plain": same,
                     "train_s": round(t_train, 3), "model_err_max": model.info()["err_max"],
                     "model_nheavy": model.info()["nheavy"],
                     "splitters_exact": L["splitters_exact"], "splitters_repaired": L["splitters_repaired"],
                     "repair_rounds": L["repair_rounds"], "plain_bisect_rounds": st_p[-1]["bisect_rounds"],
                     "t_findp_ms": {"learned": round(L["t_findp_ms"], 4), "plain": round(st_p[-1]["t_findp_ms"], 4)},
                     "buckets_overflow": L["buckets_overflow"],
                     "buckets_overflow_first_sort": L["buckets_overflow_first"],
                     "plain_sides": L["plain_sides"],
                     "buckets_lds_radix": L["buckets_lds_radix"],
                     "buckets_listed": L["buckets_listed"],
                     "phases_ms": {k2: round(L[k2], 4) for k2 in PHASES},
                     "kernels_ms": dict(zip(LEARNED_KERNELS, [round(x, 4) for x in L["pass_ms"]])),
                     "plain_stages_ms": plain_stages(st_p[-1]),
                     "plain_buckets_overflow": st_p[-1]["buckets_overflow"],
                     "plain_buckets_lds_radix": st_p[-1]["buckets_lds_radix"]}
        model.close()
        log("%s: %s" % (dist, json.dumps(out[dist])))
        if not (ok and same):
            print("bench.py: %s output check failed" % dist, file=sys.stderr)
            return 6
    worst = min(v["learned_vs_plain"] for v in out.values())
    result = {"metric": "Mrecords/s, learned path on config-3 key distributions (2^%d x 16 B, %d virtual ranks)"
                        % (args.log2n, P),
              "value": round(min(v["learned_Mrecords_s"] for v in out.values()), 1), "unit": "Mrecords/s",
              "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
              "config": {"workload": "config 3 distributions " + ",".join(out), "records": n, "vranks": P},
              "worst_learned_vs_plain": worst, "distributions": out}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
