#!/usr/bin/env python3
"""Benchmark of the MI355X PIR server answer path.

Workload (the configuration the headline metric is quoted on): ring degree
2048, plaintext modulus t = 32771 (PlainMod 16, the reference's setting for
every N = 2048 preset; --plain-bits 17 gives t = 65537), l = 5, l_key = 10; first dimension 512 x 1024 further
(524,288 plaintexts, 8 GiB NTT-form database resident in HBM); a batch of 64
queries for distinct seeded indices, each from its own client with its own
keys.  A step is one batch answered end to end (expand, convert, first
dimension, other dimensions, modulus switch).  Every response is checked: the
client must decrypt the exact seeded record.

Also reported: the single-query path (queries/s, p50/p99 latency, database
GB/s scanned), the first-dimension kernel against the HBM roofline (HIP
events on the stream the kernel runs on), and the CPU oracle timed on this
host on a bounded sample, whose responses must equal the GPU's bit for bit.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PIR queries/s + DB GB/s scanned, 1-query vs batched-64 (bit-exact); p50/p99"
HBM_PEAK_GBPS = 8000.0
N = 2048


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, p):
    return float(np.percentile(np.asarray(xs), p))


def measured_traffic(kernel: str, fst: int, other: int):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this kernel and shape (profiles/), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{fst}x{other}_hbm_bytes.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1]), os.path.relpath(path, ROOT)
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fst", type=int, default=512)
    ap.add_argument("--other", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--single-reps", type=int, default=20)
    ap.add_argument("--db-seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-queries", type=int, default=2)
    ap.add_argument("--plain-bits", type=int, default=16)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (profiler byte-counter scale)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "the database would shard by rows with one 32 KiB gather per GPU, see DESIGN.md", file=sys.stderr)
        return 2

    from onionpirv2_amd import GpuServer, dbgen
    from onionpirv2_amd.client import GpuClient

    fst, other, B = args.fst, args.other, args.batch
    srv = GpuServer(fst, other, plain_bits=args.plain_bits)
    info = srv.info
    num_pt = info["num_pt"]
    db_bytes = num_pt * N * 8
    log(f"[bench] shape {fst}x{other}: {num_pt} plaintexts, NTT-form DB {db_bytes / 2**30:.1f} GiB")
    t0 = time.time()
    srv.db_generate(args.db_seed, 0)
    log(f"[bench] device DB ingest {time.time() - t0:.2f} s")
    if args.calibrate_copy:
        n = srv.tap_copy(args.calibrate_copy << 20)
        log(f"[bench] calibration copy k_copy: {n} bytes read, {n} bytes written")

    rng = np.random.default_rng(12345)
    idx = rng.choice(num_pt, size=B, replace=False)
    clients, queries = [], []
    t0 = time.time()
    for i in range(B):
        c = GpuClient(fst, other, client_id=i, seed=1000 + i, plain_bits=args.plain_bits)
        srv.set_galois_key(i, c.galois_keys())
        srv.set_rgsw_key(i, c.rgsw_key())
        clients.append(c)
        queries.append(c.query(int(idx[i]), nonce=i))
    queries = np.stack(queries)
    ids = np.arange(B, dtype=np.uint64)
    log(f"[bench] {B} clients keyed, queries built in {time.time() - t0:.2f} s")

    # ---- single-query path ---------------------------------------------------------
    single_lat, cores, stages = [], [], []
    srv.answer(0, queries[0])  # warm
    for r in range(args.single_reps):
        i = r % B
        t = time.perf_counter()
        srv.answer(i, queries[i])
        single_lat.append((time.perf_counter() - t) * 1e3)
        st = srv.stage_times()
        cores.append(st["first_dim_core"])
        stages.append(st)
    log(f"[bench] single query p50 {pct(single_lat, 50):.3f} ms")

    # ---- batched path (the timed steps) -----------------------------------------------
    for _ in range(args.warmup):
        srv.answer_batch(ids, queries)
    step_ms, bstages = [], []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        resps = srv.answer_batch(ids, queries)
        step_ms.append((time.perf_counter() - t) * 1e3)
        bstages.append(srv.stage_times())
    total_s = time.perf_counter() - t_all

    # ---- correctness: every client decrypts its exact record ----------------------------
    ok = 0
    for i in range(B):
        pt = clients[i].decrypt(resps[i])
        ok += int(np.array_equal(pt, dbgen.plaintext(args.db_seed, 0, int(idx[i]), t=info["t"])))
    bit_exact = ok == B
    log(f"[bench] records recovered {ok}/{B}")

    # ---- roofline of the first-dimension kernel ------------------------------------------
    core_ms = float(np.mean(cores))
    alg_bytes = db_bytes + N * fst * 16 + other * 2 * N * 8
    achieved = alg_bytes / (core_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_first_dim_b1", fst, other)

    out = {
        "metric": METRIC,
        "value": B * args.steps / total_s,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded records uniform in [0,t), seeded clients and indices",
        "config": {"workload": f"{fst}x{other} plaintexts (N=2048, t={info['t']}, l=5, l_key=10), "
                               f"{db_bytes / 2**30:.0f} GiB NTT-form DB in HBM, batch of {B} queries",
                   "fst_dim": fst, "other_dim": other, "batch": B, "poly_degree": N,
                   "plain_modulus": info["t"]},
        "batched": {"queries_per_s": B * args.steps / total_s, "step_ms": float(np.mean(step_ms)),
                    "p50_ms": pct(step_ms, 50), "p99_ms": pct(step_ms, 99),
                    "db_GBps_scanned": db_bytes * B / (np.mean(step_ms) * 1e-3) / 1e9,
                    "stage_ms": {k: float(np.mean([s[k] for s in bstages])) for k in bstages[0]}
                    if bstages else {}},
        "single_query": {"queries_per_s": 1e3 / float(np.mean(single_lat)),
                         "p50_ms": pct(single_lat, 50), "p99_ms": pct(single_lat, 99),
                         "db_GBps_scanned": db_bytes / (float(np.mean(single_lat)) * 1e-3) / 1e9,
                         "stage_ms": {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}},
        "batched_vs_single": (B * args.steps / total_s) / (1e3 / float(np.mean(single_lat))),
        "bit_exact": bit_exact,
        "roofline": {"kernel": "k_first_dim_b1", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic, "traffic_source": traffic_src, "avg_ms": core_ms,
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": None,
    }

    # ---- CPU baseline: the oracle on this host, same inputs, bounded sample -------------------
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, clients, queries, idx, srv)
    print(json.dumps(out), flush=True)
    return 0 if bit_exact else 1


def cpu_baseline(args, clients, queries, idx, srv):
    """The CPU oracle (oracle/, a C restatement of the reference server, one
    core) on the same database and queries; its responses must equal the GPU's."""
    import concurrent.futures as cf

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, ptr

    o = Oracle(fst=args.fst, other=args.other, plain_bits=args.plain_bits)
    t0 = time.time()
    db = np.zeros(o.num_pt * N, np.uint64)
    workers = min(16, os.cpu_count() or 1)
    chunk = (o.num_pt + workers - 1) // workers
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda w: o.L.orc_db_build_range(o.c, args.db_seed, 0, w * chunk,
                                                     min(o.num_pt, (w + 1) * chunk), ptr(db)),
                    range(workers)))
    log(f"[bench] oracle DB built in {time.time() - t0:.1f} s ({workers} threads)")
    times, match = [], 0
    for i in range(args.cpu_queries):
        c = clients[i]
        gk, rkn = c.galois_keys(), o.rgsw_key_ntt(c.rgsw_key())
        resp, tm = o.answer(db, gk, rkn, queries[i])
        times.append(tm["total"])
        match += int(np.array_equal(resp, srv.answer(i, queries[i])))
        log(f"[bench] oracle query {i}: {tm['total']:.2f} s "
            + " ".join(f"{k}={v:.2f}" for k, v in tm.items()))
    return {"value": 1.0 / float(np.mean(times)), "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_queries} full single queries of the same workload (same DB, clients, "
                      f"queries) through the C oracle, 1 thread",
            "seconds_per_query": float(np.mean(times)),
            "responses_equal_gpu": f"{match}/{args.cpu_queries}"}


if __name__ == "__main__":
    sys.exit(main())
