"""Secondary benchmark workloads of bench.py (--workload large | batch |
serve | sketch | reads | genomes).

Each returns one JSON-ready dict.  Inputs of the distance workloads are
seeded synthetic sketches from the library's tree generator, resident in HBM
before timing starts; the genomes workload starts from sequence files.

large   32,768 samples, 8 k-mer lengths, 10,000 bins: one self pass over
        536,854,528 pairs, dense core/accessory streamed to page-locked host
        memory in row blocks, plus kNN = 10 lists.
batch   256 query-vs-ref jobs of uneven shape (queries log-uniform in
        [1, 2048], references uniform in [500, 20000], 4-14 k-mer lengths
        within k = 13..39, 1,000-20,000 bins), no random-match correction
        (the equal-base form is undefined for k >= 33).  Jobs use prefixes of
        8 resident databases, one per (k-set, sketch size) class.  Timed as
        one persistent launch (graph replay, autotuned tile shape) and as the
        plain path: one ppsk_dense call per job.
serve   a resident database (default 200,000 samples, 8 k, 10,000 bins,
        about 28 GB) serving query batches of log-uniform size in [1, 512]
        with kNN = 10 and threshold 0.05 (core), equal-base correction.
        Saturation throughput (every request queued at once), then Poisson
        arrivals at a fixed rate for latency percentiles, and the plain path
        (each request alone: one query-vs-ref ppsk_dense into device memory).
sketch  the same 64 genomes held in memory, sketched by ppsk_sketch (8 k,
        10,000 bins): the GPU sketcher alone, inputs from host memory.
reads   one read set (150 bp reads at 30x of a 2 Mb genome, 1% errors)
        sketched with the count filter (min_count 3, count-min sketch).
genomes 64 synthetic 2 Mb genomes down a random tree, written as FASTA:
        constructDatabase (read, sketch on the GPU at 8 k-mer lengths and
        10,000 bins, write HDF5) and queryDatabase over the 2,016 pairs.
"""
import json
import math
import os
import threading
import time

import numpy as np

import ppsk_hip as H

KL8 = list(range(15, 30, 2))
S64 = 156


def _ops_per_pair(n_k, s64):
    return n_k * s64 * 30


# ---------------------------------------------------------------------------
def large(a):
    n = 32768
    t0 = time.time()
    raw, le, _ = H.synth(n, KL8, S64, seed=3)
    print(f"# generated {n} sketches in {time.time() - t0:.1f}s", flush=True)
    pairs = n * (n - 1) // 2
    with H.SketchSet(raw, KL8, S64, lengths=le) as s, \
            H.PinnedBuffer((pairs, 2), np.float32) as pin:
        H.self_pass(s, pin.array, kNN=10, random_mode=H.RANDOM_EQUAL_BASE)  # warm-up
        walls, kms = [], []
        for _ in range(max(1, a.steps // 5)):
            t = time.perf_counter()
            _, _, nf = H.self_pass(s, pin.array, kNN=10, random_mode=H.RANDOM_EQUAL_BASE)
            walls.append(time.perf_counter() - t)
            kms.append(H.last_timing()[0])
            if nf:
                raise RuntimeError(f"{nf} pairs failed the k-mer regression")
    wall = min(walls)
    ms = min(kms)
    return {
        "metric": "distances/sec, 32,768-sample all-vs-all streamed to host + kNN=10",
        "value": round(pairs / wall, 1), "unit": "distances/s", "n_gpus": 1,
        "wall_s": round(wall, 4), "dist_kernel_ms": round(ms, 2),
        "valu_frac": round(pairs * _ops_per_pair(8, S64) / (ms * 1e-3) / 39.3216e12, 4),
        "dense_bytes_to_host": pairs * 8,
        "host_rate_GBps": round(pairs * 8 / wall / 1e9, 2),
        "config": {"workload": "large all-vs-all, dense FP32 to pinned host + kNN=10",
                   "n_samples": n, "pairs": pairs, "kmers": KL8, "bins": S64 * 64},
    }


# ---------------------------------------------------------------------------
def _batch_jobs(n_jobs, rng):
    jobs = []
    for _ in range(n_jobs):
        q = int(math.floor(2048.0 ** rng.uniform()))
        q = min(max(q, 1), 2048)
        r = int(rng.integers(500, 20001))
        jobs.append((q, r))
    return jobs


def batch(a):
    rng = np.random.default_rng(11)
    n_classes = 8
    classes = []
    for c in range(n_classes):
        k = int(rng.integers(4, 15))
        kl = sorted(rng.choice(np.arange(13, 40), size=k, replace=False).tolist())
        s64 = int(round(rng.uniform(1000, 20000) / 64))
        classes.append((kl, s64))
    shapes = _batch_jobs(a.batch_jobs, rng)
    per_class = [[] for _ in range(n_classes)]
    for x, (q, r) in enumerate(shapes):
        per_class[x % n_classes].append((x, q, r))
    t0 = time.time()
    dbs = []
    for c, (kl, s64) in enumerate(classes):
        rmax = max(r for _, _, r in per_class[c])
        qmax = max(q for _, q, _ in per_class[c])
        raw, _, _ = H.synth(rmax + qmax, kl, s64, seed=500 + c)
        ref = H.SketchSet(raw[:rmax], kl, s64)
        qry = H.SketchSet(raw[rmax:], kl, s64)
        del raw
        dbs.append((ref, qry))
    print(f"# generated and uploaded {n_classes} databases in {time.time() - t0:.1f}s",
          flush=True)
    jobs, views, pairs, ops = [], [], 0, 0
    for c in range(n_classes):
        kl, s64 = classes[c]
        ref, qry = dbs[c]
        for x, q, r in per_class[c]:
            rv, qv = ref.prefix(r), qry.prefix(q)
            views.append((rv, qv))
            jobs.append(dict(ref=rv, query=qv, random_mode=H.RANDOM_NONE))
            pairs += q * r
            ops += q * r * _ops_per_pair(len(kl), s64)
    b = H.Batch(jobs)
    t = time.perf_counter()
    b.run()  # first run: the autotuner times each tile shape on this batch
    tune_s = time.perf_counter() - t
    chosen, per_plan, row_tiles = b.tuning()
    times, kms = [], []
    for _ in range(max(1, a.steps // 2)):
        t = time.perf_counter()
        nf, used, ms = b.run()
        times.append(time.perf_counter() - t)
        kms.append(ms)
    batch_s = min(times)
    if nf.any():
        raise RuntimeError(f"{int(nf.sum())} pairs failed the k-mer regression")
    # spot check: jobs of the batch equal their single ppsk_dense calls (the
    # largest sketch, the most k-mer lengths, the most queries, and two more)
    def pick(key):
        return max(range(len(jobs)), key=key)
    check = sorted({pick(lambda x: views[x][0].s64), pick(lambda x: len(views[x][0].klist)),
                    pick(lambda x: views[x][1].n), 0, len(jobs) // 2})
    for x in check:
        rv, qv = views[x]
        want, _, _ = H.dense(rv, qv)
        if not np.array_equal(b.result(x).view(np.uint8), want.view(np.uint8)):
            raise RuntimeError(f"batch job {x} differs from its single ppsk_dense call")
    # plain path: one ppsk_dense call per job, output in device memory
    plain_s = 0.0
    for rv, qv in views:
        d = H.device_alloc(qv.n * rv.n * 8)
        t = time.perf_counter()
        H.dense_device(rv, qv, dptr=d)
        plain_s += time.perf_counter() - t
        H.device_free(d)
    b.free()
    for rv, qv in views:
        rv.free()
        qv.free()
    for ref, qry in dbs:
        ref.free()
        qry.free()
    return {
        "metric": "distances/sec, 256 uneven query-vs-ref jobs in one persistent launch",
        "value": round(pairs / batch_s, 1), "unit": "distances/s", "n_gpus": 1,
        "batch_s": round(batch_s, 4), "batch_kernel_ms": round(min(kms), 2),
        "plain_value": round(pairs / plain_s, 1), "plain_s": round(plain_s, 4),
        "speedup_vs_plain": round(plain_s / batch_s, 2),
        "valu_frac": round(ops / (min(kms) * 1e-3) / 39.3216e12, 4),
        "autotune": {"chosen_plan": chosen, "first_run_s": round(tune_s, 3),
                     "plan_ms": {str(v): round(float(per_plan[v]), 2) for v in (2, 4, 5, 6)
                                 if per_plan[v] > 0},
                     "row_tiles_per_height": {str(h): int(row_tiles[:, x].sum())
                                              for x, h in enumerate((64, 32, 16))}},
        "checked_jobs": check,
        "config": {"workload": "batched uneven query-vs-ref jobs", "jobs": len(jobs),
                   "pairs": pairs, "queries": "log-uniform 1..2048",
                   "refs": "uniform 500..20000", "classes": [
                       {"kmers": kl, "bins": s64 * 64} for kl, s64 in classes],
                   "random_correction": "none"},
    }


# ---------------------------------------------------------------------------
def serve(a):
    n_ref = a.serve_refs
    n_pool = 4096
    t0 = time.time()
    # a wider tree than the all-vs-all workloads (core divergence up to 15%
    # root to leaf) so that the 0.05 threshold keeps close relatives only,
    # as in a real database
    raw, le, _ = H.synth(n_ref + n_pool, KL8, S64, seed=21, dcore_max=0.15)
    q_raw, q_le = raw[n_ref:].copy(), le[n_ref:].copy()
    ref = H.SketchSet(raw[:n_ref], KL8, S64, lengths=le[:n_ref])
    del raw
    print(f"# generated and uploaded {n_ref} reference sketches in {time.time() - t0:.1f}s",
          flush=True)
    rng = np.random.default_rng(4)

    def draw(m):
        out = []
        for _ in range(m):
            size = min(512, max(1, int(math.floor(512.0 ** rng.uniform()))))
            s = int(rng.integers(0, n_pool - size + 1))
            out.append((s, s + size))
        return out

    # unbatched: the same server, one request at a time (submit, wait)
    solo_req = draw(32)
    with H.Server(ref, kNN=10, cutoff=0.05, random_mode=H.RANDOM_EQUAL_BASE,
                  max_rows=512, n_slots=1) as solo:
        t = time.perf_counter()
        for s_, e_ in solo_req:
            solo.wait(solo.submit(q_raw[s_:e_], q_le[s_:e_]))
        solo_s = time.perf_counter() - t
    solo_rate = sum(e_ - s_ for s_, e_ in solo_req) * n_ref / solo_s

    # plain path: each request alone, one query-vs-ref dense launch
    plain_req = draw(48)
    plain_pairs, plain_s = 0, 0.0
    for s_, e_ in plain_req:
        with H.SketchSet(q_raw[s_:e_], KL8, S64, lengths=q_le[s_:e_]) as q:
            d = H.device_alloc((e_ - s_) * n_ref * 8)
            t = time.perf_counter()
            H.dense_device(ref, q, random_mode=H.RANDOM_EQUAL_BASE, dptr=d)
            plain_s += time.perf_counter() - t
            H.device_free(d)
        plain_pairs += (e_ - s_) * n_ref
    plain_rate = plain_pairs / plain_s

    srv = H.Server(ref, kNN=10, cutoff=0.05, random_mode=H.RANDOM_EQUAL_BASE,
                   max_rows=a.serve_max_rows, n_slots=a.serve_slots)
    try:
        # saturation: every request queued at once
        sat_req = draw(64)
        t = time.perf_counter()
        tickets = [srv.submit(q_raw[s_:e_], q_le[s_:e_]) for s_, e_ in sat_req]
        sat_res = [srv.wait(tk) for tk in tickets]
        thr_per_query = sum(len(r["thr_j"]) for r in sat_res) / sum(
            e_ - s_ for s_, e_ in sat_req)
        sat_s = time.perf_counter() - t
        sat_pairs = sum(e_ - s_ for s_, e_ in sat_req) * n_ref
        sat_rate = sat_pairs / sat_s
        sat_qps = len(sat_req) / sat_s
        # Poisson arrivals at a fixed rate
        qps = a.serve_qps or 0.7 * sat_qps
        n_req = max(8, int(qps * a.serve_seconds))
        reqs = draw(n_req)
        gaps = rng.exponential(1.0 / qps, size=n_req)
        tickets = [None] * n_req
        results = [None] * n_req

        def consumer():
            for x in range(n_req):
                while tickets[x] is None:
                    time.sleep(0.0005)
                results[x] = srv.wait(tickets[x])

        th = threading.Thread(target=consumer)
        th.start()
        t_begin = time.perf_counter()
        nxt = t_begin
        for x, (s_, e_) in enumerate(reqs):
            nxt += gaps[x]
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[x] = srv.submit(q_raw[s_:e_], q_le[s_:e_])
        th.join()
        wall = time.perf_counter() - t_begin
        st = srv.stats()
    finally:
        srv.free()
        ref.free()
    lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in results])
    pairs = sum(e_ - s_ for s_, e_ in reqs) * n_ref
    return {
        "metric": "served distances/sec with p50/p99 request latency (kNN=10, threshold 0.05)",
        "value": round(pairs / wall, 1), "unit": "distances/s", "n_gpus": 1,
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "qps": round(qps, 3), "requests": n_req, "wall_s": round(wall, 3),
        "saturation_value": round(sat_rate, 1), "saturation_qps": round(sat_qps, 3),
        "plain_value": round(plain_rate, 1),
        "speedup_vs_plain": round(sat_rate / plain_rate, 2),
        "unbatched_value": round(solo_rate, 1),
        "speedup_vs_unbatched": round(sat_rate / solo_rate, 2),
        "threshold_hits_per_query": round(thr_per_query, 1),
        "n_fail": int(sum(r["n_fail"] for r in results)),
        "server": st,
        "config": {"workload": "serving, resident reference database", "n_refs": n_ref,
                   "kmers": KL8, "bins": S64 * 64, "request_sizes": "log-uniform 1..512",
                   "kNN": 10, "threshold": 0.05, "slots": a.serve_slots, "max_rows_per_step": a.serve_max_rows,
                   "random_correction": "equal-base"},
    }


# ---------------------------------------------------------------------------
def _poisson(srv, reqs, q_raw, q_le, qps, rng):
    """Submits reqs at Poisson arrivals of rate qps while a consumer thread
    collects the results; returns (results, wall seconds)."""
    n_req = len(reqs)
    gaps = rng.exponential(1.0 / qps, size=n_req)
    tickets = [None] * n_req
    results = [None] * n_req

    def consumer():
        for x in range(n_req):
            while tickets[x] is None:
                time.sleep(0.0005)
            results[x] = srv.wait(tickets[x])

    th = threading.Thread(target=consumer)
    th.start()
    t_begin = time.perf_counter()
    nxt = t_begin
    for x, (s_, e_) in enumerate(reqs):
        nxt += gaps[x]
        dt = nxt - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        tickets[x] = srv.submit(q_raw[s_:e_], q_le[s_:e_])
    th.join()
    return results, time.perf_counter() - t_begin


def serve_headline(n_ref, seconds, max_rows, slots):
    """The serving figure of the headline line: a resident database of n_ref
    sketches (8 k-mer lengths, 10,000 bins), query batches of log-uniform
    size in [1, 512], kNN = 10 and threshold 0.05 (core), equal-base
    correction.  Saturation throughput (64 requests queued at once), then
    Poisson arrivals at 70% of that request rate for `seconds`: p50 / p99 of
    submit-to-result latency."""
    n_pool = 4096
    t0 = time.time()
    raw, le, _ = H.synth(n_ref + n_pool, KL8, S64, seed=21, dcore_max=0.15)
    q_raw, q_le = raw[n_ref:].copy(), le[n_ref:].copy()
    ref = H.SketchSet(raw[:n_ref], KL8, S64, lengths=le[:n_ref])
    del raw
    db_bytes = n_ref * len(KL8) * S64 * 14 * 8
    setup_s = time.time() - t0
    print(f"# serving: {n_ref} reference sketches generated and resident in {setup_s:.1f}s",
          flush=True)
    rng = np.random.default_rng(4)

    def draw(m):
        out = []
        for _ in range(m):
            size = min(512, max(1, int(math.floor(512.0 ** rng.uniform()))))
            s = int(rng.integers(0, n_pool - size + 1))
            out.append((s, s + size))
        return out

    srv = H.Server(ref, kNN=10, cutoff=0.05, random_mode=H.RANDOM_EQUAL_BASE,
                   max_rows=max_rows, n_slots=slots)
    try:
        srv.wait(srv.submit(q_raw[:8], q_le[:8]))  # warm-up
        sat_req = draw(64)
        st0 = srv.stats()
        t = time.perf_counter()
        tickets = [srv.submit(q_raw[s_:e_], q_le[s_:e_]) for s_, e_ in sat_req]
        for tk in tickets:
            srv.wait(tk)
        sat_s = time.perf_counter() - t
        sat_steps = srv.stats()["steps"] - st0["steps"]
        sat_queries = sum(e_ - s_ for s_, e_ in sat_req)
        sat_rate = sat_queries * n_ref / sat_s
        sat_qps = len(sat_req) / sat_s
        qps = 0.7 * sat_qps
        reqs = draw(max(8, int(qps * seconds)))
        results, wall = _poisson(srv, reqs, q_raw, q_le, qps, rng)
    finally:
        srv.free()
        ref.free()
    lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in results])
    pairs = sum(e_ - s_ for s_, e_ in reqs) * n_ref
    return {
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "value": round(pairs / wall, 1), "unit": "distances/s",
        "saturation_value": round(sat_rate, 1),
        # int32 matching work and reference-set passes at saturation: each
        # step streams the resident set at least once (once per 64-row tile)
        "saturation_int_ops_per_s": round(sat_rate * _ops_per_pair(len(KL8), S64), 1),
        "saturation_valu_frac": round(sat_rate * _ops_per_pair(len(KL8), S64) / 39.3216e12, 4),
        "saturation_set_bytes_per_s": round(sat_steps * db_bytes / sat_s, 1),
        "saturation_steps": sat_steps,
        "load": "Poisson arrivals at 70% of the saturation request rate",
        "qps": round(qps, 3), "requests": len(reqs), "window_s": round(wall, 2),
        "n_fail": int(sum(r["n_fail"] for r in results)),
        "n_refs": n_ref, "kmers": KL8, "bins": S64 * 64,
        "request_sizes": "log-uniform 1..512 sketches", "kNN": 10, "threshold": 0.05,
        "slots": slots, "max_rows_per_step": max_rows, "random_correction": "equal-base",
    }


# ---------------------------------------------------------------------------
def _tree_genomes(n_leaves, length, rng, n_blocks=400, block=1000):
    """Leaves of a random binary tree over one i.i.d. ACGT ancestor (codes
    0-3).  Each branch applies SNPs at a rate drawn from [0.02%, 0.8%], so a
    leaf lies 0.1-5% from the root; the accessory pool holds n_blocks 1 kb
    blocks, each branch flipping each block's presence with probability
    [0, 5%], so leaves carry 0-20% accessory sequence."""
    anc = rng.integers(0, 4, size=length, dtype=np.uint8)
    pool = rng.integers(0, 4, size=(n_blocks, block), dtype=np.uint8)
    leaves = []

    def grow(core, present, k):
        if k == 1:
            leaves.append(np.concatenate([core, pool[present].reshape(-1)]))
            return
        left = int(rng.integers(1, k))
        for part in (left, k - left):
            c = core.copy()
            m = rng.random(length) < rng.uniform(0.0002, 0.008)
            c[m] = (c[m] + rng.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) % 4
            p = present ^ (rng.random(n_blocks) < rng.uniform(0.0, 0.05))
            grow(c, p, part)

    grow(anc, rng.random(n_blocks) < 0.5, n_leaves)
    return leaves


def genomes(a):
    """64 synthetic 2 Mb genomes -> FASTA -> constructDatabase (GPU sketching)
    -> queryDatabase all-vs-all (2,016 pairs)."""
    import shutil
    import tempfile

    import pp_sketchlib as pp

    n, length = 64, 2_000_000
    rng = np.random.default_rng(17)
    t0 = time.time()
    g = _tree_genomes(n, length, rng)
    work = tempfile.mkdtemp(prefix="ppsk_genomes_")
    try:
        letters = np.frombuffer(b"ACGT", np.uint8)
        files, names = [], []
        for i, codes in enumerate(g):
            p = os.path.join(work, f"g{i}.fa")
            with open(p, "wb") as f:
                f.write(b">contig1\n" + letters[codes].tobytes() + b"\n")
            files.append([p])
            names.append(f"genome_{i}")
        bases = int(sum(x.size for x in g))
        print(f"# wrote {n} genomes ({bases / 1e6:.1f} Mb) in {time.time() - t0:.1f}s", flush=True)
        db = os.path.join(work, "db")
        walls, kms, qwalls = [], [], []
        for rep in range(max(2, a.steps // 5 + 1)):
            if os.path.exists(db + ".h5"):
                os.remove(db + ".h5")
            t = time.perf_counter()
            pp.constructDatabase(db, names, files, KL8, S64, calc_random=False, num_threads=16)
            walls.append(time.perf_counter() - t)
            kms.append(H.last_timing()[0])
            t = time.perf_counter()
            d = pp.queryDatabase(db, db, names, names, KL8)
            qwalls.append(time.perf_counter() - t)
            print(f"# rep {rep}: construct {walls[-1]:.2f}s (sketch kernel {kms[-1]:.1f} ms), "
                  f"query {qwalls[-1] * 1e3:.1f} ms", flush=True)
        assert d.shape == (n * (n - 1) // 2, 2)
        # Monte Carlo random-match tables on the last database, then the
        # query again with them
        t = time.perf_counter()
        pp.addRandom(db, names, KL8, True, 16)
        t_random = time.perf_counter() - t
        t = time.perf_counter()
        d_mc = pp.queryDatabase(db, db, names, names, KL8)
        t_query_mc = time.perf_counter() - t
        print(f"# random tables {t_random:.2f}s, query with them {t_query_mc * 1e3:.1f} ms",
              flush=True)
        kmers = sum((x.size - k + 1) for x in g for k in KL8)
        wall, ms, qwall = min(walls[1:]), min(kms[1:]), min(qwalls[1:])
        line = {
            "metric": "genomes sketched/s (8 k, 10,000 bins) via constructDatabase, "
                      "then all-vs-all queryDatabase",
            "value": round(n / wall, 2), "unit": "genomes/s", "n_gpus": 1,
            "construct_wall_s": round(wall, 3),
            "sketch_kernel_ms": round(ms, 2),
            "kmer_hashes_per_s": round(kmers / (ms * 1e-3), 1),
            "query_wall_ms": round(qwall * 1e3, 2),
            "end_to_end_s": round(wall + qwall, 3),
            "random_tables_s": round(t_random, 3),
            "query_mc_wall_ms": round(t_query_mc * 1e3, 2),
            "query_mc_finite": bool(np.isfinite(d_mc).all()),
            "config": {"workload": "64 tree genomes of 2 Mb (+0-20% accessory), FASTA files",
                       "bases": bases, "kmers": KL8, "bins": S64 * 64, "pairs": d.shape[0]},
        }
        if not a.no_cpu:
            line["cpu_baseline"] = _genomes_cpu_baseline(g, names, db, d, kmers)
        return line
    finally:
        shutil.rmtree(work, ignore_errors=True)


def _genomes_cpu_baseline(g, names, db, d, n_kmers):
    """The oracle's sketch (one thread) of the first genome at every k, scaled
    to all 64; its all-vs-all on the stored sketches (16 threads)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import oracle as O
    import pp_sketchlib as pp

    seq0 = np.frombuffer(b"ACGT", np.uint8)[g[0]].tobytes().decode()
    t = time.perf_counter()
    for k in KL8:
        words, _, _ = O.sketch([seq0], k, S64, True)
        assert np.array_equal(words, pp.getSketch(db, names[0])["sketch"][k])
    dt = time.perf_counter() - t
    per_genome = dt * (n_kmers / len(g)) / sum(g[0].size - k + 1 for k in KL8)
    raw = np.array([[pp.getSketch(db, nm)["sketch"][k] for k in KL8] for nm in names], np.uint64)
    lens = np.array([pp.getSketch(db, nm)["length"] for nm in names], np.uint64)
    t = time.perf_counter()
    want, _, _, _ = O.query_db(raw, KL8, S64, 14, lens, random=O.Random(1, True), num_threads=16)
    dq = time.perf_counter() - t
    return {"value": round(1.0 / per_genome, 3), "unit": "genomes/s", "cores": 1, "kind": "port",
            "sample": f"genome 0 sketched at 8 k ({dt:.1f} s, words equal to the GPU's); "
                      f"query of 2,016 pairs on 16 threads {dq * 1e3:.1f} ms",
            "query_max_abs_diff": float(np.abs(d - want).max())}


def sketch(a):
    """ppsk_sketch over 64 tree genomes of 2 Mb held in host memory."""
    n, length = 64, 2_000_000
    g = _tree_genomes(n, length, np.random.default_rng(17))
    letters = np.frombuffer(b"ACGT", np.uint8)
    samples = [[letters[x].tobytes()] for x in g]
    kmers = sum(x.size - k + 1 for x in g for k in KL8)
    bases = int(sum(x.size for x in g))
    print(f"# {n} genomes, {bases / 1e6:.1f} Mb, {kmers / 1e9:.3f} G k-mers over 8 k", flush=True)
    walls, kms = [], []
    for rep in range(max(2, a.steps)):
        t = time.perf_counter()
        H.sketch(samples, KL8, S64)
        walls.append(time.perf_counter() - t)
        kms.append(H.last_timing()[0])
        print(f"# rep {rep}: wall {walls[-1] * 1e3:.1f} ms, kernels {kms[-1]:.2f} ms", flush=True)
    ms, wall = min(kms[1:]), min(walls[1:])
    return {
        "metric": "k-mer hashes/s, GPU sketching (8 k, 10,000 bins, canonical)",
        "value": round(kmers / (ms * 1e-3), 1), "unit": "k-mers/s", "n_gpus": 1,
        "kernels_ms": round(ms, 3), "wall_ms": round(wall * 1e3, 2),
        "genomes_per_s_wall": round(n / wall, 1),
        "config": {"workload": "64 tree genomes of 2 Mb (+0-20% accessory) in host memory",
                   "bases": bases, "kmers": KL8, "bins": S64 * 64, "kmer_windows": kmers},
    }


def reads(a):
    """ppsk_sketch_filtered over one read set: 150 bp reads at 30x of a 2 Mb
    genome (1% substitution errors), min_count 3, count-min sketch."""
    rng = np.random.default_rng(23)
    length, read_len, cover = 2_000_000, 150, 30
    genome = rng.integers(0, 4, size=length, dtype=np.uint8)
    n_reads = length * cover // read_len
    starts = rng.integers(0, length - read_len + 1, size=n_reads)
    idx = starts[:, None] + np.arange(read_len)[None, :]
    codes = genome[idx]
    err = rng.random(codes.shape) < 0.01
    codes[err] = (codes[err] + rng.integers(1, 4, size=int(err.sum()), dtype=np.uint8)) % 4
    letters = np.frombuffer(b"ACGT", np.uint8)
    flat = letters[codes].tobytes()
    contigs = [flat[i * read_len:(i + 1) * read_len] for i in range(n_reads)]
    windows = sum(n_reads * (read_len - k + 1) for k in KL8)
    print(f"# {n_reads} reads, {windows / 1e9:.3f} G windows over 8 k", flush=True)
    walls, kms = [], []
    for rep in range(max(2, a.steps)):
        t = time.perf_counter()
        H.sketch([contigs], KL8, S64, min_count=3)
        walls.append(time.perf_counter() - t)
        kms.append(H.last_timing()[0])
        print(f"# rep {rep}: wall {walls[-1] * 1e3:.1f} ms, device {kms[-1]:.2f} ms", flush=True)
    ms, wall = min(kms[1:]), min(walls[1:])
    line = {
        "metric": "read k-mer windows/s, GPU sketching with the count filter (min_count 3)",
        "value": round(windows / (ms * 1e-3), 1), "unit": "windows/s", "n_gpus": 1,
        "device_ms": round(ms, 2), "wall_ms": round(wall * 1e3, 2),
        "config": {"workload": "150 bp reads, 30x of a 2 Mb genome, 1% errors, count-min",
                   "reads": n_reads, "kmers": KL8, "bins": S64 * 64, "windows": windows},
    }
    if not a.no_cpu:
        import sys
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(
            os.path.abspath(__file__))), "oracle"))
        import oracle as O
        sub = contigs[: n_reads // 20]
        t = time.perf_counter()
        O.sketch_filtered(sub, 15, S64, True, 3, False)
        dt = time.perf_counter() - t
        w_sub = len(sub) * (read_len - 15 + 1)
        line["cpu_baseline"] = {"value": round(w_sub / dt, 1), "unit": "windows/s", "cores": 1,
                                "kind": "port",
                                "sample": f"the oracle's filtered sketch of 1/20 of the reads at "
                                          f"k = 15 ({dt:.1f} s)"}
    return line
