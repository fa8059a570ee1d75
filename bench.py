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
GB/s scanned; eager launches and the replayed hipGraph), the batched
first-dimension kernel against the int8 matrix-core roofline and the
single-query one against the HBM roofline (HIP events on the stream the
kernels run on), two batches in flight on two lanes, the serving loop under Poisson
arrivals (p50/p99 from submit to response), three databases of different
shape resident together (128x256, 256x1024, 512x4096) answering bursts of
7 / 33 / 100 queries, the same plus 512x8192 (100.5 GiB resident) with 1000
clients' keys in one shared device key store under Poisson arrivals (50 to
2000 queries/s), a 256x512 single-query shape, the batch at t = 65537, and
the CPU oracle timed on this host on a bounded sample (3 untimed and 10 timed
queries on one core, built -march=native here, then one oracle server per
usable CPU up to 16), whose responses must equal the GPU's bit for bit.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
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
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 matrix-core peak
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


def mfma_kernel(tuning: dict, batch: int, fst: int):
    """(name, file key) of the matrix-core first-dimension kernel the server
    launches for a batch under this tuning."""
    cb = 1 if batch <= 16 else 2 if batch <= 32 else 4
    lds = tuning.get("fd_mfma_variant") == 1
    if tuning.get("db_planes") and tuning.get("fd_mfma_variant") == 2 and fst % 64 == 0:
        return f"k_first_dim_mfma_tiles<{cb}>", "k_first_dim_mfma_tiles"
    if tuning.get("db_planes"):
        return (f"k_first_dim_mfma_lds<{cb}, true>", "k_first_dim_mfma_lds_pre") if lds else \
               (f"k_first_dim_mfma_pre<{cb}>", "k_first_dim_mfma_pre")
    return (f"k_first_dim_mfma_lds<{cb}, false>", "k_first_dim_mfma_lds") if lds else \
           (f"k_first_dim_mfma<{cb}>", "k_first_dim_mfma")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fst", type=int, default=512)
    ap.add_argument("--other", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--single-reps", type=int, default=200,
                    help="single queries per leg (tuned, graph, eager): enough for a p99")
    ap.add_argument("--db-seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-queries", type=int, default=10, help="measured oracle queries (1 core)")
    ap.add_argument("--cpu-warmup", type=int, default=3, help="untimed oracle queries before them")
    ap.add_argument("--plain-bits", type=int, default=16)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="threads of the multi-core CPU baseline (one oracle server each); 0 = the CPUs this "
                         "process may run on, at most 16 (a GPU box's share); 1 = skip")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the built-in launch shapes instead of tuning on the live database")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the two-lane, serving and multi-database measurements")
    ap.add_argument("--serve-rates", default="100,500,1000,2000", help="Poisson arrival rates (queries/s)")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--config5-clients", type=int, default=1000,
                    help="clients of the four-database serving workload (0 = skip it)")
    ap.add_argument("--config5-rates", default="50,200,500,1000,2000", help="its Poisson rates (queries/s)")
    ap.add_argument("--config5-workers", type=int, default=2,
                    help="batches in flight over the four databases (the serving loop's worker pool)")
    ap.add_argument("--config5-lanes", type=int, default=1, help="lanes per database in the serving loop")
    ap.add_argument("--config5-only", action="store_true",
                    help="run only the four-database serving workload and print its results as the JSON line")
    ap.add_argument("--no-single-legs", dest="single_legs", action="store_false",
                    help="skip the single-query legs that force graph replay and eager launches")
    ap.add_argument("--multi-db-linger-us", type=int, default=0,
                    help="how long the three-database bursts' first arrivals wait for the rest of their burst")
    ap.add_argument("--no-t65537", dest="t65537", action="store_false",
                    help="skip the batch-64 line at t = 65537 (PlainMod 17)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (profiler byte-counter scale)")
    ap.add_argument("--valu-probe", type=int, default=0, metavar="ROUNDS",
                    help="also run the VALU issue-rate probe (ROUNDS radix-8 butterfly passes per thread) "
                         "at 1..4 waves per SIMD: the ceiling the transform-bound kernels' counters are read against")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); "
              "the database would shard by rows, each GPU answering its rows and one 32 KiB gather "
              "per GPU combining them", file=sys.stderr)
        return 2

    if args.config5_only:
        c5 = config5(args)
        print(json.dumps({"config5": c5}), flush=True)
        return 0 if c5["bit_exact"] else 1

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

    # 100 = config 4's largest burst: more than one pass, so the batch-split
    # choice (64 + 36 or 50 + 50) is timed too
    tune_sizes = [1, B] + ([100] if B == 64 else [])
    tuning = srv.autotune(tune_sizes, reps=3) if args.autotune else srv.get_tuning()
    log(f"[bench] tuning {tuning}")

    # ---- single-query path ---------------------------------------------------------
    single_lat, cores, stages = [], [], []
    single_resp = {}
    srv.answer(0, queries[0])  # warm
    for r in range(args.single_reps):
        i = r % B
        t = time.perf_counter()
        single_resp[i] = srv.answer(i, queries[i])
        single_lat.append((time.perf_counter() - t) * 1e3)
        st = srv.stage_times()
        cores.append(st["first_dim_core"])
        stages.append(st)
    log(f"[bench] single query p50 {pct(single_lat, 50):.3f} ms")
    graph_single = eager_single = None
    if args.single_legs:
        graph_single = single_query_graph(srv, queries, args.single_reps, single_resp)
        eager_single = single_query_graph(srv, queries, args.single_reps, single_resp, use_graph=0)

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

    # ---- rooflines of the first-dimension kernels ----------------------------------------
    # per-kernel and per-stage events exist only for eagerly launched
    # pipelines: where the tuning replays graphs (use_graph 1: every size, 2:
    # single queries), time the kernels and stages on extra eager passes (same
    # kernels, same shapes); the device totals stay the replayed calls' own
    dev_ms = [st["total"] for st in bstages]
    dev_single_ms = [st["total"] for st in stages]
    single_graphed = tuning.get("use_graph") in (1, 2)
    batch_graphed = tuning.get("use_graph") == 1
    if single_graphed or batch_graphed:
        srv.set_tuning(use_graph=0)
        if single_graphed:
            stages = []
            for r in range(max(3, args.single_reps // 2)):
                srv.answer(r % B, queries[r % B])
                stages.append(srv.stage_times())
            cores = [st["first_dim_core"] for st in stages]
        if batch_graphed:
            bstages = []
            for _ in range(max(1, args.steps // 2)):
                srv.answer_batch(ids, queries)
                bstages.append(srv.stage_times())
        srv.set_tuning(**tuning)
    fd_cores = [st["first_dim_core"] for st in bstages]
    # single query: k_first_dim_b1 streams the database once (HBM bound)
    core_ms = float(np.mean(cores))
    alg_bytes = db_bytes + N * fst * 16 + other * 2 * N * 8
    achieved = alg_bytes / (core_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_first_dim_b1", fst, other)
    # batch: k_first_dim_mfma, 64 signed-digit int8 products per 64x64-bit
    # product, 2 * B * num_pt * N such products per launch (matrix-core bound)
    mfma_ms = float(np.mean(fd_cores))
    ncp = 32 if B <= 16 else 64 if B <= 32 else 128
    int8_ops = 2 * 64 * (ncp * num_pt * N)  # the padded column count is what the tiles compute
    useful_ops = 2 * 64 * (2 * B * num_pt * N)
    mfma_alg_bytes = db_bytes + N * fst * ncp * 8 + N * other * 2 * B * 8
    mfma_name, mfma_key = mfma_kernel(tuning, B, fst)
    mfma_traffic, mfma_traffic_src = measured_traffic(f"{mfma_key}_b{B}", fst, other)

    # value: device time of each step (HIP events from the first kernel to the
    # last, queries already in HBM); the wall clock around answer_batch also
    # holds the PCIe copies of queries and responses and is reported apart
    value = B * 1e3 / float(np.mean(dev_ms))
    single_qps_dev = 1e3 / float(np.mean(dev_single_ms))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(dev_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded records uniform in [0,t), seeded clients and indices",
        "config": {"workload": f"{fst}x{other} plaintexts (N=2048, t={info['t']}, l=5, l_key=10), "
                               f"{db_bytes / 2**30:.0f} GiB NTT-form DB in HBM, batch of {B} queries",
                   "fst_dim": fst, "other_dim": other, "batch": B, "poly_degree": N,
                   "plain_modulus": info["t"]},
        "batched": {"queries_per_s": value, "step_ms": float(np.mean(dev_ms)),
                    "p50_ms": pct(dev_ms, 50), "p99_ms": pct(dev_ms, 99),
                    "db_GBps_scanned": db_bytes * B / (np.mean(dev_ms) * 1e-3) / 1e9,
                    "pcie_inclusive_queries_per_s": B * args.steps / total_s,
                    "pcie_inclusive_step_ms": float(np.mean(step_ms)),
                    "stage_ms": {k: float(np.mean([s[k] for s in bstages])) for k in bstages[0]}
                    if bstages else {}},
        "single_query": {"queries_per_s": single_qps_dev, "device_ms": float(np.mean(dev_single_ms)),
                         "p50_ms": pct(single_lat, 50), "p99_ms": pct(single_lat, 99),
                         "latency_note": "p50/p99: wall clock of one synchronous call, PCIe copies included",
                         "pcie_inclusive_queries_per_s": 1e3 / float(np.mean(single_lat)),
                         "db_GBps_scanned": db_bytes / (float(np.mean(dev_single_ms)) * 1e-3) / 1e9,
                         "stage_ms": {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}},
        "batched_vs_single": value / single_qps_dev,
        "bit_exact": bit_exact,
        "roofline": {"kernel": mfma_name, "bound": "mfma", "achieved": useful_ops / (mfma_ms * 1e-3) / 1e12,
                     "peak": INT8_PEAK_TOPS, "unit": "TOP/s (int8)",
                     "frac": useful_ops / (mfma_ms * 1e-3) / 1e12 / INT8_PEAK_TOPS,
                     "traffic": mfma_traffic, "traffic_source": mfma_traffic_src, "avg_ms": mfma_ms,
                     "algorithmic_ops": useful_ops, "issued_ops": int8_ops,
                     "algorithmic_bytes": mfma_alg_bytes,
                     "hbm_GBps": mfma_alg_bytes / (mfma_ms * 1e-3) / 1e9},
        "roofline_single_query": {"kernel": "k_first_dim_b1", "bound": "hbm", "achieved": achieved,
                                  "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                                  "traffic": traffic, "traffic_source": traffic_src, "avg_ms": core_ms,
                                  "algorithmic_bytes": alg_bytes},
        "cpu_baseline": None,
        "tuning": tuning,
        "autotune_sizes": tune_sizes if args.autotune else None,
    }

    out["single_query_graph"] = graph_single
    out["single_query_eager"] = eager_single
    if args.valu_probe:
        # blocks = 256 w: w workgroups per compute unit, one wave per SIMD each
        out["valu_probe"] = [{"waves_per_simd": w, "blocks": 256 * w, "rounds": args.valu_probe,
                              "ms": srv.tap_valu_probe(256 * w, args.valu_probe),
                              "butterflies": 256 * w * 256 * 12 * args.valu_probe} for w in (1, 2, 3, 4)]
        log(f"[bench] VALU probe ms by waves per SIMD: {[round(p['ms'], 3) for p in out['valu_probe']]}")
    ok_all = bit_exact and all(leg["bit_exact"] for leg in (graph_single, eager_single) if leg)


    # ---- two lanes: two batches in flight on two streams --------------------------------
    if not args.no_extras:
        lanes = two_lanes(srv, ids, queries, resps, args.steps)
        out["batched_two_lanes"] = lanes
        ok_all = ok_all and lanes["bit_exact"]
        for i in range(B):
            if i not in single_resp:
                single_resp[i] = srv.answer(i, queries[i])
        srv_rates = [float(r) for r in args.serve_rates.split(",") if r]
        out["serving"] = serving(srv, queries, single_resp, srv_rates, args.serve_seconds)
        ok_all = ok_all and all(r["bit_exact"] for r in out["serving"]["rates"])
        del srv, clients  # free the 8 GiB database before the other workloads
        out["config_256x512"] = single_shape(args, 256, 512)
        ok_all = ok_all and out["config_256x512"]["bit_exact"]
        if args.t65537:
            out["batched_t65537"] = batched_t65537(args, tuning)
        out["multi_db"] = multi_db(args)
        ok_all = ok_all and out["multi_db"]["bit_exact"]
        if args.config5_clients:
            out["config5"] = config5(args)
            ok_all = ok_all and out["config5"]["bit_exact"]

    # ---- CPU baseline: the oracle on this host, same inputs, bounded sample -------------------
    if not args.no_cpu:
        if args.no_extras:
            out["cpu_baseline"] = cpu_baseline(args, clients, queries, idx, srv)
        else:
            out["cpu_baseline"] = cpu_baseline_fresh(args, queries, idx, single_resp)
    print(json.dumps(out), flush=True)
    return 0 if ok_all else 1


def single_query_graph(srv, queries, reps, single_resp, use_graph=1):
    """The single-query pipeline captured once as a hipGraph and replayed
    (use_graph=1), or launched eagerly kernel by kernel (use_graph=0)."""
    saved = srv.get_tuning()
    srv.set_tuning(use_graph=use_graph)
    try:
        srv.answer(0, queries[0])  # capture
        lat, same = [], True
        for r in range(reps):
            i = r % len(queries)
            t = time.perf_counter()
            got = srv.answer(i, queries[i])
            lat.append((time.perf_counter() - t) * 1e3)
            if i in single_resp:
                same = same and np.array_equal(got, single_resp[i])
    finally:
        srv.set_tuning(**saved)
    log(f"[bench] single query ({'graph' if use_graph else 'eager'}) p50 {pct(lat, 50):.3f} ms")
    return {"queries_per_s": 1e3 / float(np.mean(lat)), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "bit_exact": bool(same)}


def two_lanes(srv, ids, queries, ref, steps):
    """Two host threads each run `steps` batches on their own lane."""
    import threading
    srv.set_lanes(2)
    srv.answer_batch_lane(1, ids, queries)  # allocate lane 1's scratch
    res, errs = [], []

    def run(lane):
        try:
            for _ in range(steps):
                r, _ = srv.answer_batch_lane(lane, ids, queries)
                res.append(np.array_equal(r, ref))
        except Exception as e:  # noqa: BLE001
            errs.append(repr(e))

    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    n = 2 * steps * len(ids)
    log(f"[bench] two lanes: {n / dt:.1f} queries/s")
    return {"queries_per_s": n / dt, "batches": 2 * steps, "batch": len(ids),
            "bit_exact": bool(not errs and all(res)), "errors": errs}


def serving(srv, queries, single_resp, rates, seconds):
    """Poisson arrivals through the serving loop (continuous batching, two
    lanes); latency is submit -> response ready.  Every response must equal
    the synchronous single-query answer to the same query."""
    from onionpirv2_amd.serve import GpuServe
    B = len(queries)
    out = []
    with GpuServe(max_batch=64, lanes_per_db=2) as sv:
        db, cap = sv.add_db(srv)
        rng = np.random.default_rng(99)
        for rate in rates:
            n = max(16, int(rate * seconds))
            gaps = rng.exponential(1.0 / rate, size=n)
            sv.reset_stats()
            tickets = []
            t0 = time.perf_counter()
            due = t0
            for k in range(n):
                due += gaps[k]
                while time.perf_counter() < due:
                    pass
                tickets.append((k % B, sv.submit(db, k % B, queries[k % B])))
            t_last = time.perf_counter()
            resps = [sv.wait(t) for i, t in tickets]
            drain_ms = (time.perf_counter() - t_last) * 1e3  # last arrival -> last response
            same = all(np.array_equal(r, single_resp[i]) for r, (i, t) in zip(resps, tickets))
            st = sv.stats(db)
            offered = n / (due - t0)
            log(f"[bench] serving at {rate:.0f} q/s offered: {st['queries_per_s']:.1f} q/s done, "
                f"p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms, drain {drain_ms:.1f} ms, "
                f"mean batch {st['mean_batch']:.1f}")
            out.append({"offered_qps": offered, "target_qps": rate, "queries": n,
                        "completed_qps": st["queries_per_s"], "drain_ms": drain_ms,
                        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                        "max_ms": st["max_ms"], "mean_batch": st["mean_batch"], "batches": st["batches"],
                        "bit_exact": bool(same and st["failed"] == 0)})
    return {"lanes": 2, "max_batch": 64, "batch_cap": cap, "rates": out}


def single_shape(args, fst, other, reps=20):
    """Single queries on another shape (256x512: 2 GiB database): latency,
    device stage times and the first-dimension kernel's HBM rate."""
    from onionpirv2_amd import GpuServer, dbgen
    from onionpirv2_amd.client import GpuClient
    s = GpuServer(fst, other, plain_bits=args.plain_bits)
    s.db_generate(args.db_seed + 7, 0)
    c = GpuClient(fst, other, client_id=1, seed=77, plain_bits=args.plain_bits)
    s.set_galois_key(1, c.galois_keys())
    s.set_rgsw_key(1, c.rgsw_key())
    num_pt = s.info["num_pt"]
    idx = [int(i) for i in np.random.default_rng(5).integers(0, num_pt, size=reps)]
    qs = [c.query(i, nonce=500 + k) for k, i in enumerate(idx)]
    s.answer(1, qs[0])
    lat, st, ok = [], [], 0
    for k in range(reps):
        t = time.perf_counter()
        r = s.answer(1, qs[k])
        lat.append((time.perf_counter() - t) * 1e3)
        st.append(s.stage_times())
        ok += int(np.array_equal(c.decrypt(r), dbgen.plaintext(args.db_seed + 7, 0, idx[k], t=s.info["t"])))
    dev = [x["total"] for x in st]
    db_bytes = num_pt * N * 8
    if s.get_tuning()["use_graph"]:
        # replayed calls report only their total: stage times from eager passes
        s.set_tuning(use_graph=0)
        st = []
        for k in range(reps // 2):
            s.answer(1, qs[k])
            st.append(s.stage_times())
    core = float(np.mean([x["first_dim_core"] for x in st]))
    alg = db_bytes + N * fst * 16 + other * 2 * N * 8
    s.close()
    log(f"[bench] {fst}x{other}: single query p50 {pct(lat, 50):.3f} ms, first dim {alg / core / 1e6:.0f} GB/s")
    return {"shape": f"{fst}x{other}", "db_GiB": db_bytes / 2**30, "queries": reps,
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "device_ms": float(np.mean(dev)),
            "stage_ms": {k: float(np.mean([x[k] for x in st])) for k in st[0]},
            "first_dim_GBps": alg / (core * 1e-3) / 1e9, "first_dim_frac_of_hbm_peak": alg / (core * 1e-3) / 1e9 / HBM_PEAK_GBPS,
            "records_ok": ok, "bit_exact": ok == reps}


def config5(args):
    """Four databases resident together (128x256, 256x1024, 512x4096,
    512x8192: 100.5 GiB NTT form) sharing one device key store that holds the
    keys of `config5_clients` clients, served by the serving loop (one lane per
    database, scratch admitted from a 120 GiB budget) under Poisson arrivals;
    each query goes to a uniformly drawn database, client and record.  Every
    response must decrypt to its record."""
    from onionpirv2_amd import GpuServer, dbgen
    from onionpirv2_amd.client import GpuClient
    from onionpirv2_amd.serve import GpuServe
    shapes = [(128, 256), (256, 1024), (512, 4096), (512, 8192)]
    nc = args.config5_clients
    t0 = time.time()
    servers = []
    for d, (fst, other) in enumerate(shapes):
        s = GpuServer(fst, other, plain_bits=args.plain_bits)
        s.db_generate(args.db_seed + 11 + d, 0)
        if servers:
            s.share_keys(servers[0])
        servers.append(s)
    resident = sum(s.info["num_pt"] * N * 8 for s in servers)
    t_db = time.time() - t0
    t0 = time.time()
    tall = shapes[-1]  # the tallest expansion: its keys serve all four databases
    for c in range(nc):
        cl = GpuClient(tall[0], tall[1], client_id=c, seed=90000 + c, plain_bits=args.plain_bits)
        servers[0].set_galois_key(c, cl.galois_keys())
        servers[0].set_rgsw_key(c, cl.rgsw_key())
    t_keys = time.time() - t0
    # device bytes: the Galois key words are held with their Shoup companions
    key_bytes = nc * (2 * servers[-1].info["galois_key_words"] + servers[-1].info["rgsw_key_words"]) * 8
    log(f"[bench] config5: {resident / 2**30:.1f} GiB resident ({t_db:.1f} s), {nc} clients' keys "
        f"({key_bytes / 2**30:.2f} GiB, one shared store) in {t_keys:.1f} s")
    rng = np.random.default_rng(555)
    qclients = {}

    def client(d, c):
        if (d, c) not in qclients:
            qclients[(d, c)] = GpuClient(shapes[d][0], shapes[d][1], client_id=c, seed=90000 + c,
                                         plain_bits=args.plain_bits)
        return qclients[(d, c)]

    rates = [float(r) for r in args.config5_rates.split(",") if r]
    runs, ok_all = [], True
    with GpuServe(max_batch=64, lanes_per_db=args.config5_lanes, scratch_budget=120 << 30,
                  workers=args.config5_workers) as sv:
        caps = [sv.add_db(s)[1] for s in servers]
        # unloaded latency: one query at a time through the loop, per database
        unloaded = []
        for d in range(4):
            lat = []
            for k in range(7):
                c = int(rng.integers(nc))
                i = int(rng.integers(servers[d].info["num_pt"]))
                q = client(d, c).query(i, nonce=50000 + 10 * d + k)
                t = time.perf_counter()
                r = sv.wait(sv.submit(d, c, q), db_id=d)
                lat.append((time.perf_counter() - t) * 1e3)
                ok_all = ok_all and np.array_equal(client(d, c).decrypt(r), dbgen.plaintext(
                    args.db_seed + 11 + d, 0, i, t=servers[d].info["t"]))
            unloaded.append(float(np.median(lat[2:])))
        log("[bench] config5 unloaded p50 per database: " +
            ", ".join(f"{a}x{b} {u:.2f} ms" for (a, b), u in zip(shapes, unloaded)))
        for rate in rates:
            n = max(40, int(rate * args.serve_seconds))
            work = []
            for k in range(n):
                d = int(rng.integers(4))
                c = int(rng.integers(nc))
                i = int(rng.integers(servers[d].info["num_pt"]))
                work.append((d, c, i, client(d, c).query(i, nonce=int(rate) * 100000 + k)))
            gaps = rng.exponential(1.0 / rate, size=n)
            sv.reset_stats()
            tickets = []
            t0 = time.perf_counter()
            due = t0
            for k, (d, c, i, q) in enumerate(work):
                due += gaps[k]
                while time.perf_counter() < due:
                    pass
                tickets.append(sv.submit(d, c, q))
            t_last = time.perf_counter()
            resps = [sv.wait(tk, db_id=d) for (d, c, i, q), tk in zip(work, tickets)]
            # from the last arrival to the last response: about one batch when the
            # loop keeps up, growing with the run's length when it does not
            drain_ms = (time.perf_counter() - t_last) * 1e3
            good = 0
            for (d, c, i, q), r in zip(work, resps):
                good += int(np.array_equal(client(d, c).decrypt(r),
                                           dbgen.plaintext(args.db_seed + 11 + d, 0, i, t=servers[d].info["t"])))
            st = sv.stats()
            per_db = []
            for d in range(4):
                sd = sv.stats(d)
                per_db.append({"shape": f"{shapes[d][0]}x{shapes[d][1]}", "queries": sd["completed"],
                               "queries_per_s": sd["queries_per_s"], "p50_ms": sd["p50_ms"], "p99_ms": sd["p99_ms"],
                               "p50_over_unloaded": sd["p50_ms"] / unloaded[d], "mean_batch": sd["mean_batch"]})
            ok_all = ok_all and good == n
            log(f"[bench] config5 at {rate:.0f} q/s: {st['queries_per_s']:.1f} q/s done, p50 {st['p50_ms']:.1f} ms "
                f"p99 {st['p99_ms']:.1f} ms, drain {drain_ms:.0f} ms, {good}/{n} records; p50 / unloaded per "
                "database " + " ".join(f"{x['p50_over_unloaded']:.1f}" for x in per_db))
            runs.append({"target_qps": rate, "offered_qps": n / (due - t0), "queries": n,
                         "completed_qps": st["queries_per_s"], "drain_ms": drain_ms,
                         "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                         "max_ms": st["max_ms"], "mean_batch": st["mean_batch"], "records_ok": good,
                         "per_db": per_db})
    for s in reversed(servers):
        s.close()
    return {"resident_GiB": resident / 2**30, "shapes": [f"{a}x{b}" for a, b in shapes], "clients": nc,
            "key_store_GiB": key_bytes / 2**30, "batch_caps": caps, "scratch_budget_GiB": 120,
            "unloaded_p50_ms": unloaded,
            "workers": args.config5_workers, "lanes_per_db": args.config5_lanes, "runs": runs,
            "bit_exact": bool(ok_all)}


def multi_db(args):
    """Three databases resident together (128x256, 256x1024, 512x4096: 0.5 +
    4 + 32 GiB), each tuned by the autotuner for the per-database shares of
    the bursts (3, 11 and 34 queries); bursts of 7 / 33 / 100 queries spread
    over them go through the serving loop; every record must decrypt."""
    from onionpirv2_amd import GpuServer, dbgen
    from onionpirv2_amd.client import GpuClient
    from onionpirv2_amd.serve import GpuServe
    shapes = [(128, 256), (256, 1024), (512, 4096)]
    n_clients = 8
    servers, clients = [], {}
    t0 = time.time()
    for d, (fst, other) in enumerate(shapes):
        s = GpuServer(fst, other, plain_bits=args.plain_bits)
        s.db_generate(args.db_seed + 1 + d, 0)
        for c in range(n_clients):
            cl = GpuClient(fst, other, client_id=c, seed=3000 + 100 * d + c, plain_bits=args.plain_bits)
            s.set_galois_key(c, cl.galois_keys())
            s.set_rgsw_key(c, cl.rgsw_key())
            clients[(d, c)] = cl
        servers.append(s)
    resident = sum(s.info["num_pt"] * N * 8 for s in servers)
    log(f"[bench] multi-db: {resident / 2**30:.1f} GiB resident, set up in {time.time() - t0:.1f} s")
    t0 = time.time()
    tunings = [s.autotune([3, 11, 34], reps=2) for s in servers]
    log(f"[bench] multi-db: tuned in {time.time() - t0:.1f} s")
    rng = np.random.default_rng(4242)
    bursts, ok = [], True
    with GpuServe(max_batch=64, lanes_per_db=1, linger_us=args.multi_db_linger_us) as sv:
        for s in servers:
            sv.add_db(s)
        for burst in (7, 33, 100):
            work = []
            for k in range(burst):
                d = k % 3
                c = int(rng.integers(n_clients))
                i = int(rng.integers(servers[d].info["num_pt"]))
                work.append((d, c, i, clients[(d, c)].query(i, nonce=burst * 1000 + k)))
            sv.reset_stats()
            t = time.perf_counter()
            tickets = [sv.submit(d, c, q) for d, c, _, q in work]
            resps = [sv.wait(tk, db_id=w[0]) for tk, w in zip(tickets, work)]
            wall = time.perf_counter() - t
            good = sum(int(np.array_equal(clients[(d, c)].decrypt(r),
                                          dbgen.plaintext(args.db_seed + 1 + d, 0, i, t=servers[d].info["t"])))
                       for (d, c, i, _), r in zip(work, resps))
            ok = ok and good == burst
            per_db = []
            for d in range(3):
                st = sv.stats(d)
                per_db.append({"shape": f"{shapes[d][0]}x{shapes[d][1]}", "queries": st["completed"],
                               "queries_per_s": st["queries_per_s"], "p50_ms": st["p50_ms"], "max_ms": st["max_ms"],
                               "batches": st["batches"], "mean_batch": st["mean_batch"]})
            log(f"[bench] multi-db burst {burst}: {wall * 1e3:.1f} ms, {good}/{burst} records")
            bursts.append({"burst": burst, "wall_ms": wall * 1e3, "queries_per_s": burst / wall,
                           "records_ok": good, "per_db": per_db})
    for s in servers:
        s.close()
    return {"resident_GiB": resident / 2**30, "shapes": [f"{a}x{b}" for a, b in shapes], "bursts": bursts,
            "tuning": {f"{a}x{b}": t for (a, b), t in zip(shapes, tunings)}, "autotune_sizes": [3, 11, 34],
            "linger_us": args.multi_db_linger_us,
            "bit_exact": bool(ok)}


def batched_t65537(args, tuning):
    """The headline batch at t = 65537 (PlainMod 17, the configurations'
    "t = 2^17"): same shape, seeds and tuning, device time per batch.  At this
    t the 512x1024 shape is at the edge of the noise budget, so a query can
    fail to decrypt (the oracle's response is the same); the count is
    reported."""
    from onionpirv2_amd import GpuServer, dbgen
    from onionpirv2_amd.client import GpuClient
    fst, other, B = args.fst, args.other, args.batch
    s = GpuServer(fst, other, plain_bits=17)
    try:
        s.db_generate(args.db_seed, 0)
        s.set_tuning(**tuning)
        num_pt, t = s.info["num_pt"], s.info["t"]
        idx = np.random.default_rng(12345).choice(num_pt, size=B, replace=False)
        clients, qs = [], []
        for i in range(B):
            c = GpuClient(fst, other, client_id=i, seed=1000 + i, plain_bits=17)
            s.set_galois_key(i, c.galois_keys())
            s.set_rgsw_key(i, c.rgsw_key())
            clients.append(c)
            qs.append(c.query(int(idx[i]), nonce=i))
        qs = np.stack(qs)
        ids = np.arange(B, dtype=np.uint64)
        for _ in range(args.warmup):
            s.answer_batch(ids, qs)
        dev = []
        for _ in range(args.steps):
            resps = s.answer_batch(ids, qs)
            dev.append(s.stage_times()["total"])
        ok = sum(int(np.array_equal(clients[i].decrypt(resps[i]), dbgen.plaintext(args.db_seed, 0, int(idx[i]), t=t)))
                 for i in range(B))
    finally:
        s.close()
    log(f"[bench] t = 65537: {B * 1e3 / np.mean(dev):.1f} queries/s, {ok}/{B} records decrypt")
    return {"plain_modulus": t, "batch": B, "queries_per_s": B * 1e3 / float(np.mean(dev)),
            "step_ms": float(np.mean(dev)), "records_ok": ok}


def host_cpu():
    """Model name and CPU counts of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_usable": usable}


def native_oracle():
    """Build the oracle for this host's own ISA (-march=native, the
    reference's flags) and point the oracle loader at it; the tests keep the
    portable build.  Returns the build's flags."""
    import subprocess
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=300)
        os.environ["ONION_ORACLE_LIB"] = os.path.join(ROOT, "oracle", "_build", "liboracle_native.so")
        return "gcc -O3 -march=native"
    except Exception as e:  # noqa: BLE001
        log(f"[bench] native oracle build failed ({e!r}); portable build used")
        return "gcc -O3 -march=x86-64-v3"


def cpu_baseline_fresh(args, queries, idx, single_resp):
    """cpu_baseline against the single-query GPU responses already taken."""
    class _Resp:
        def answer(self, i, q):
            return single_resp[i]
    from onionpirv2_amd.client import GpuClient
    clients = [GpuClient(args.fst, args.other, client_id=i, seed=1000 + i, plain_bits=args.plain_bits)
               for i in range(args.cpu_warmup + args.cpu_queries)]
    return cpu_baseline(args, clients, queries, idx, _Resp())


def cpu_baseline(args, clients, queries, idx, srv):
    """The CPU oracle (oracle/, a C restatement of the reference server, one
    core, built with -march=native on this host) on the same database and
    queries: cpu_warmup untimed queries, then cpu_queries timed ones (the
    reference's protocol, src/tests.cpp:16); every response must equal the
    GPU's."""
    import concurrent.futures as cf

    flags = native_oracle()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, ptr

    o = Oracle(fst=args.fst, other=args.other, plain_bits=args.plain_bits)
    t0 = time.time()
    db = np.zeros(o.num_pt * N, np.uint64)
    workers = max(1, min(16, host_cpu()["cpus_usable"]))
    chunk = (o.num_pt + workers - 1) // workers
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda w: o.L.orc_db_build_range(o.c, args.db_seed, 0, w * chunk,
                                                     min(o.num_pt, (w + 1) * chunk), ptr(db)),
                    range(workers)))
    log(f"[bench] oracle DB built in {time.time() - t0:.1f} s ({workers} threads)")
    times, match = [], 0
    n = args.cpu_warmup + args.cpu_queries
    for i in range(n):
        c = clients[i]
        gk, rkn = c.galois_keys(), o.rgsw_key_ntt(c.rgsw_key())
        resp, tm = o.answer(db, gk, rkn, queries[i])
        if i >= args.cpu_warmup:
            times.append(tm["total"])
        match += int(np.array_equal(resp, srv.answer(i, queries[i])))
        log(f"[bench] oracle query {i}{' (warm-up)' if i < args.cpu_warmup else ''}: {tm['total']:.2f} s "
            + " ".join(f"{k}={v:.2f}" for k, v in tm.items()))
    out = {"value": 1.0 / float(np.mean(times)), "unit": "queries/s", "cores": 1, "kind": "port",
           "sample": f"{args.cpu_warmup} untimed + {args.cpu_queries} timed full single queries of the same "
                     f"workload (same DB, clients, queries) through the C oracle, 1 thread",
           "seconds_per_query": float(np.mean(times)), "p50_s": pct(times, 50), "p99_s": pct(times, 99),
           "build": flags, **host_cpu(),
           "responses_equal_gpu": f"{match}/{n}"}
    if args.cpu_cores != 1:
        out["multicore"] = cpu_multicore(args, o, db, clients, queries, srv)
    return out


def cpu_multicore(args, o, db, clients, queries, srv):
    """k independent single-threaded oracle servers (one per host core, the
    database shared read-only), each answering one full query at once."""
    import concurrent.futures as cf

    from oracle_lib import Oracle

    k = args.cpu_cores if args.cpu_cores > 1 else max(1, min(16, host_cpu()["cpus_usable"]))
    n = len(clients)
    orcs = [Oracle(fst=args.fst, other=args.other, plain_bits=args.plain_bits) for _ in range(k)]
    keys = [(c.galois_keys(), o.rgsw_key_ntt(c.rgsw_key())) for c in clients]
    t0 = time.time()
    with cf.ThreadPoolExecutor(k) as ex:
        resps = list(ex.map(lambda i: orcs[i].answer(db, keys[i % n][0], keys[i % n][1], queries[i % n])[0],
                            range(k)))
    wall = time.time() - t0
    match = sum(int(np.array_equal(r, srv.answer(i % n, queries[i % n]))) for i, r in enumerate(resps))
    log(f"[bench] oracle on {k} cores: {k} queries in {wall:.2f} s")
    return {"value": k / wall, "unit": "queries/s", "cores": k, "kind": "port",
            "sample": f"{k} full single queries of the same workload, one per thread, each on its own "
                      f"oracle server over the shared database (threads: the CPUs this process may use, at "
                      f"most 16, a GPU box's CPU share)",
            **host_cpu(), "responses_equal_gpu": f"{match}/{k}"}


if __name__ == "__main__":
    sys.exit(main())
