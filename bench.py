#!/usr/bin/env python3
"""Benchmark of the core/accessory distance path on one MI355X.

Workload (one step): all-vs-all core/accessory distances of 8,192 synthetic
sketches, 8 k-mer lengths (k = 15, 17, ..., 29), 10,000 bins (156 words x 14
bit layers per k), equal-base random-match correction, dense FP32 output
(33,550,336 pairs) written to HBM.  Sketches are generated with a seeded
tree-evolution model and resident in HBM before timing starts.

Prints one JSON line:
  value      distances (pairs) per second over the timed steps
  roofline   int32 VALU issue bound of the distance-tile kernel: algorithmic
             ops per pair = sum_k s64 * (2 * bbits + 2) = 37,440, divided by
             the kernel's HIP-event duration on its launch stream
  cpu_baseline  the CPU oracle (C++/OpenMP restatement of the reference CPU
             path), compiled on this host with -march=native, on a bounded
             sample of the same workload: median of 3 runs at 16 threads and
             of 3 runs on every core of the process's CPU set
  serving    the metric's latency figure: a resident database of 200,000
             sketches (same k-mer lengths and sketch size) serving Poisson
             query batches of 1-512 sketches at 70% of the measured saturation
             rate with kNN = 10 and threshold 0.05: p50 / p99 request latency
             and the saturation throughput
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pp-sketchlib_amd"))

N_SAMPLES = 8192
KLIST = list(range(15, 30, 2))
S64 = 156
BBITS = 14
OPS_PER_PAIR = len(KLIST) * S64 * (2 * BBITS + 2)  # 37,440
# vendor peak: 256 CUs x 64 lanes x 2.4 GHz int32 VALU ops
PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# published GPU figure for this metric (RTX 2080 Ti, 8 k, 10k bins; README)
PUBLISHED_GPU = 6.0e6
METRIC = "distances/sec (8 k, 10k bins) on 1 MI355X; int-VALU %peak; p99 assign ms"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_SAMPLES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="samples of the CPU baseline's all-vs-all sample")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="CPU baseline runs per thread count (the median is reported)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving measurement of the headline line")
    ap.add_argument("--headline-serve-seconds", type=float, default=20.0,
                    help="Poisson arrival window of the headline line's serving figure")
    ap.add_argument("--copy-only", type=int, default=0, metavar="BYTES",
                    help="run one plain copy kernel of BYTES and exit (HBM counter calibration)")
    ap.add_argument("--workload", default="plain", choices=["plain", "large", "batch", "serve", "sketch", "reads", "genomes"],
                    help="plain: 8,192-sample all-vs-all (the headline line); large: "
                         "32,768-sample self pass streamed to host + kNN; batch: 256 uneven "
                         "query-vs-ref jobs; serve: resident 200,000-sample database under "
                         "Poisson query batches; sketch: GPU sketching of 64 synthetic 2 Mb "
                         "genomes; reads: count-filtered sketching of a 30x read set; "
                         "genomes: the 64 genomes through constructDatabase + queryDatabase")
    ap.add_argument("--serve-refs", type=int, default=200_000)
    ap.add_argument("--serve-qps", type=float, default=0.0,
                    help="Poisson arrival rate (0: 70%% of the measured saturation rate)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-max-rows", type=int, default=128,
                    help="query rows per serving step (larger requests are split)")
    ap.add_argument("--serve-slots", type=int, default=2,
                    help="serving steps in flight (worker streams)")
    ap.add_argument("--batch-jobs", type=int, default=256)
    return ap.parse_args()


def _bench_lib():
    import ctypes
    L = ctypes.CDLL(os.path.join(ROOT, "pp-sketchlib_amd", "libppsk_bench.so"))
    L.ppsk_bench_copy.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)]
    L.ppsk_bench_valu_peak.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    return L


def copy_only(nbytes):
    """One copy of nbytes (read nbytes, write nbytes) for counter calibration."""
    import ctypes
    ms = ctypes.c_double()
    rc = _bench_lib().ppsk_bench_copy(0, nbytes, 1, ctypes.byref(ms))
    if rc:
        raise RuntimeError("copy kernel failed")
    print(json.dumps({"copy_bytes": nbytes, "ms": ms.value,
                      "GB_per_s": 2 * nbytes / ms.value / 1e6}), flush=True)


def measured_valu_peak():
    """Chained v_bitop3_b32 microbenchmark: int32 VALU ops/s this GPU sustains."""
    import ctypes
    ops, ms = ctypes.c_double(), ctypes.c_double()
    if _bench_lib().ppsk_bench_valu_peak(0, 16384, 2000, ctypes.byref(ops), ctypes.byref(ms)):
        return None
    return ops.value / 1e12


def traffic_from_profile():
    """HBM bytes per distance pass from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "dist_batch_allvsall_8192_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_pass")


def _host_cpu():
    """CPU model, SMT state and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = "on" if f.read().strip() == "1" else "off"
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, smt, cores


def _native_oracle():
    """Builds the oracle for this host's CPU (-march=native, the reference's
    own Makefile flag) into a temporary directory and points the oracle
    module at it; returns the compile flags used."""
    import subprocess
    import tempfile
    out = tempfile.mkdtemp(prefix="ppsk_oracle_")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                    f"NATIVE_DIR={out}"], check=True, stdout=subprocess.DEVNULL)
    os.environ["PPSK_ORACLE_LIB"] = os.path.join(out, "liboracle_native.so")
    return "-O3 -march=native -ffp-contract=off -fno-fast-math -fopenmp"


def cpu_baseline(raw, lengths, n_sample, runs):
    """Median distances/s of the oracle's all-vs-all over the first n_sample
    samples: `runs` runs at 16 threads and `runs` on every available core."""
    import statistics
    try:
        flags = _native_oracle()
    except Exception as e:  # no compiler: the portable build (x86-64-v3)
        print(f"# native oracle build failed ({e}); using oracle/liboracle.so", file=sys.stderr)
        flags = "-O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math -fopenmp"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    sub = raw[:n_sample]
    le = lengths[:n_sample]
    pairs = n_sample * (n_sample - 1) // 2
    model, smt, avail = _host_cpu()

    def rate(threads):
        t0 = time.perf_counter()
        O.query_db(sub, KLIST, S64, BBITS, le, random=O.Random(1, True), num_threads=threads)
        return pairs / (time.perf_counter() - t0)

    rate(16)  # first touch: library load, page faults
    r16 = [rate(16) for _ in range(runs)]
    rall = [rate(avail) for _ in range(runs)] if avail != 16 else r16
    med16 = statistics.median(r16)
    print(f"# cpu baseline: {med16:.3e} distances/s at 16 threads, "
          f"{statistics.median(rall):.3e} at {avail}", file=sys.stderr, flush=True)
    return {
        "value": med16,
        "unit": "distances/s",
        "cores": 16,
        "kind": "port",
        "sample": f"all-vs-all of the first {n_sample} of the {N_SAMPLES} samples "
                  f"({pairs} pairs per run), same k-mer lengths / sketch size / random "
                  f"correction; median of {runs} runs",
        "runs_16_threads": [round(x, 1) for x in r16],
        "all_cores": {"threads": avail, "value": round(statistics.median(rall), 1),
                      "runs": [round(x, 1) for x in rall]},
        "cpu_model": model,
        "smt": smt,
        "build": flags,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(the distance path shards by tile blocks across GPUs; only one GPU is measured here)", file=sys.stderr)
        sys.exit(2)
    if a.copy_only:
        copy_only(a.copy_only)
        return
    if a.workload != "plain":
        import bench_workloads as W
        line = {"large": W.large, "batch": W.batch, "serve": W.serve,
                "sketch": W.sketch, "reads": W.reads, "genomes": W.genomes}[a.workload](a)
        print(json.dumps(line), flush=True)
        return
    import ppsk_hip as H

    t0 = time.time()
    raw, lengths, _ = H.synth(a.n, KLIST, S64, seed=1)
    print(f"# generated {a.n} sketches in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    H.device_init(0)
    ref = H.SketchSet(raw, KLIST, S64, lengths=lengths)
    pairs = a.n * (a.n - 1) // 2
    dout = H.device_alloc(pairs * 8)
    for _ in range(a.warmup):
        H.dense_device(ref, random_mode=H.RANDOM_EQUAL_BASE, dptr=dout)
    kernel_ms = []
    launches = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, nf = H.dense_device(ref, random_mode=H.RANDOM_EQUAL_BASE, dptr=dout)
        ms, nl, _ = H.last_timing()
        kernel_ms.append(ms)
        launches += nl
        if nf:
            raise RuntimeError(f"{nf} pairs failed the k-mer regression")
    wall = time.perf_counter() - t0
    print(f"# {a.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)

    value = pairs * a.steps / wall
    # one step is one pass over every pair: the persistent tile kernel
    # (dist_batch) plus, when its last round of whole tiles would be mostly
    # idle, one launch of the per-(tile, k) kernel for those tiles; the HIP
    # events bracket both launches of the pass on their stream
    pass_ms = sum(kernel_ms) / a.steps
    achieved = OPS_PER_PAIR * pairs / (pass_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": round(PEAK_TOPS, 3),
        "unit": "TOP/s",
        "frac": round(achieved / PEAK_TOPS, 4),
        "traffic": traffic_from_profile(),
        "kernel": "dist_batch + dist_tiles_diagonal (last partial round: the diagonal tiles)",
        "kernel_ms": round(pass_ms, 3),
        "launches_per_step": round(launches / a.steps, 2),
        "ops_per_pair": OPS_PER_PAIR,
        "peak_measured": None,
    }
    peak_meas = measured_valu_peak()
    if peak_meas:
        roofline["peak_measured"] = round(peak_meas, 3)
    H.device_free(dout)
    ref.free()
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(raw, lengths, min(a.cpu_sample, a.n), a.cpu_runs)
    del raw
    serving = None
    if not a.no_serve:
        import bench_workloads as W
        serving = W.serve_headline(a.serve_refs, a.headline_serve_seconds,
                                   a.serve_max_rows, a.serve_slots)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "distances/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED_GPU, 2),
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "all-vs-all core/accessory distances, dense FP32 to HBM",
            "n_samples": a.n,
            "pairs": pairs,
            "kmers": KLIST,
            "bins": S64 * 64,
            "bbits": BBITS,
            "random_correction": "equal-base",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
