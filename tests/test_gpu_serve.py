"""Concurrent lanes, captured pipelines and the serving loop on the GPU.

Every response produced through a lane, a replayed hipGraph or the serving
loop must equal, bit for bit, the synchronous single-query answer of the same
server on the same client and query (itself pinned against the CPU oracle in
test_gpu_parity.py), and must decrypt to the seeded record.  One response per
database is also checked against the oracle directly.
"""
import threading

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N = 2048
SHAPES = [(64, 16), (128, 32)]  # two databases of different shape served together
N_CLIENTS = 4
PLAIN_BITS = 16


@pytest.fixture(scope="module")
def dbs():
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    servers, clients = [], {}
    for d, (fst, other) in enumerate(SHAPES):
        srv = GpuServer(fst, other, plain_bits=PLAIN_BITS)
        srv.db_generate(100 + d, 0)
        for cid in range(N_CLIENTS):
            c = GpuClient(fst, other, client_id=cid, seed=500 + cid, plain_bits=PLAIN_BITS)
            srv.set_galois_key(cid, c.galois_keys())
            srv.set_rgsw_key(cid, c.rgsw_key())
            clients[(d, cid)] = c
        servers.append(srv)
    yield dict(servers=servers, clients=clients)
    for s in servers:
        s.close()


def record(d, idx, srv):
    from onionpirv2_amd import dbgen
    return dbgen.plaintext(100 + d, 0, idx, t=srv.info["t"])


def make_queries(dbs, d, n, seed):
    srv = dbs["servers"][d]
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, srv.info["num_pt"], size=n)
    cids = rng.integers(0, N_CLIENTS, size=n)
    qs = np.stack([dbs["clients"][(d, int(c))].query(int(i), nonce=seed * 1000 + k)
                   for k, (c, i) in enumerate(zip(cids, idx))])
    return cids.astype(np.uint64), idx, qs


def test_oracle_anchor(dbs):
    """One synchronous answer per database equals the oracle's."""
    for d, (fst, other) in enumerate(SHAPES):
        srv = dbs["servers"][d]
        o = Oracle(fst=fst, other=other, plain_bits=PLAIN_BITS)
        db = o.db_build(100 + d, 0)
        c = dbs["clients"][(d, 1)]
        q = c.query(5, nonce=1)
        exp, _ = o.answer(db, c.galois_keys(), o.rgsw_key_ntt(c.rgsw_key()), q)
        assert np.array_equal(srv.answer(1, q), exp)


def test_lanes_match_lane0(dbs):
    srv = dbs["servers"][1]
    srv.set_lanes(3)
    cids, idx, qs = make_queries(dbs, 1, 6, seed=1)
    ref = srv.answer_batch(cids, qs)
    for lane in (0, 1, 2):
        got, t = srv.answer_batch_lane(lane, cids, qs)
        assert np.array_equal(got, ref), f"lane {lane}"
        assert t["total"] > 0 and t["first_dim_core"] > 0
    with pytest.raises(ValueError):
        srv.answer_batch_lane(3, cids, qs)
    for i in range(len(cids)):
        assert np.array_equal(ref[i], srv.answer(int(cids[i]), qs[i]))


def test_lanes_concurrent_threads(dbs):
    """Three host threads drive three lanes at once, three batches each."""
    srv = dbs["servers"][1]
    srv.set_lanes(3)
    work = [make_queries(dbs, 1, n, seed=10 + k) for k, n in enumerate((1, 5, 9))]
    refs = [srv.answer_batch(c, q) for c, _, q in work]
    results, errors = {}, []

    def run(lane):
        try:
            for rep in range(3):
                c, _, q = work[lane]
                results[(lane, rep)] = srv.answer_batch_lane(lane, c, q)[0]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (lane, rep), got in results.items():
        assert np.array_equal(got, refs[lane]), f"lane {lane} rep {rep}"


def test_graph_replays_match_eager(dbs):
    """The captured pipeline is replayed several times per batch size, with
    eager work and synchronisations of the program between replays; every
    replay equals the eager result for its own queries."""
    srv = dbs["servers"][0]
    sets = {n: [make_queries(dbs, 0, n, seed=100 * n + r) for r in range(4)] for n in (1, 5)}
    saved = srv.get_tuning()
    assert saved["use_graph"] == 2  # the default: single queries replay a graph
    try:
        srv.set_tuning(use_graph=0)
        eager = {n: [srv.answer_batch(c, q) for c, _, q in sets[n]] for n in sets}
        for mode in (1, 2):  # every size replays / single queries replay, batches run eagerly
            srv.set_tuning(use_graph=mode)
            for r in range(4):
                for n in (1, 5):
                    c, idx, q = sets[n][r]
                    got = srv.answer_batch(c, q)
                    assert np.array_equal(got, eager[n][r]), f"use_graph {mode} replay {r} batch {n}"
                    # eager work between replays: an NTT round trip on the same lane
                    x = np.arange(2 * N, dtype=np.uint64)
                    assert np.array_equal(srv.tap_ntt(srv.tap_ntt(x), inverse=True), x)
            c, idx, q = sets[1][0]
            assert np.array_equal(srv.answer(int(c[0]), q[0]), eager[1][0][0])
        t = srv.stage_times()
        assert t["total"] > 0
        for i in range(5):
            c, idx, q = sets[5][3]
            pt = dbs["clients"][(0, int(c[i]))].decrypt(got[i])
            assert np.array_equal(pt, record(0, int(idx[i]), srv))
    finally:
        srv.set_tuning(**saved)


def test_serve_mixed_databases_and_clients(dbs):
    """Queries for two databases and four clients, interleaved, submitted from
    two host threads: each response equals the query answered alone."""
    from onionpirv2_amd.serve import GpuServe
    servers = dbs["servers"]
    work = []
    for d in (0, 1):
        c, idx, q = make_queries(dbs, d, 24, seed=700 + d)
        work += [(d, int(c[i]), int(idx[i]), q[i]) for i in range(24)]
    rng = np.random.default_rng(3)
    order = rng.permutation(len(work))
    with GpuServe(max_batch=8, lanes_per_db=2) as sv:
        ids = [sv.add_db(s)[0] for s in servers]
        assert ids == [0, 1]
        tickets = [None] * len(work)

        def submit(part):
            for k in part:
                d, cid, _, q = work[k]
                tickets[k] = sv.submit(d, cid, q)

        halves = [order[0::2], order[1::2]]
        threads = [threading.Thread(target=submit, args=(h,)) for h in halves]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        got = [sv.wait(tickets[k], db_id=work[k][0]) for k in range(len(work))]
        st = sv.stats()
        assert st["completed"] == len(work) and st["failed"] == 0
        assert st["batches"] >= len(work) / 8
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        for d in (0, 1):
            sd = sv.stats(d)
            assert sd["completed"] == 24 and sd["batch_cap"] == 8
    for k, (d, cid, idx, q) in enumerate(work):
        srv = servers[d]
        assert np.array_equal(got[k], srv.answer(cid, q)), f"query {k}"
        pt = dbs["clients"][(d, cid)].decrypt(got[k])
        assert np.array_equal(pt, record(d, idx, srv)), f"query {k}"


def test_serve_unknown_client_fails_alone(dbs):
    from onionpirv2_amd import OnionKeyError
    from onionpirv2_amd.serve import GpuServe
    srv = dbs["servers"][0]
    c, idx, q = make_queries(dbs, 0, 6, seed=900)
    with GpuServe(max_batch=8, lanes_per_db=1, linger_us=20000) as sv:
        sv.add_db(srv)
        tickets = [sv.submit(0, int(c[i]) if i != 3 else 99, q[i]) for i in range(6)]
        for i, t in enumerate(tickets):
            if i == 3:
                with pytest.raises(OnionKeyError):
                    sv.wait(t)
            else:
                assert np.array_equal(sv.wait(t), srv.answer(int(c[i]), q[i]))
        assert sv.stats()["failed"] == 1


def test_serve_budget_admission(dbs):
    from onionpirv2_amd import OnionError
    from onionpirv2_amd.serve import GpuServe
    s0, s1 = dbs["servers"]
    b0, b1 = s0.scratch_bytes_per_query(), s1.scratch_bytes_per_query()
    assert b1 > b0 > 0
    assert s0.scratch_bytes(64) == 64 * b0
    assert s0.scratch_bytes(1) < s0.scratch_bytes(3) < s0.scratch_bytes(4)
    # two lanes of three queries on db 0, then less than one query's lanes of db 1
    budget = 2 * s0.scratch_bytes(3) + s0.scratch_bytes(1)
    assert 2 * s0.scratch_bytes(4) > budget and 2 * s1.scratch_bytes(1) > s0.scratch_bytes(1)
    with GpuServe(max_batch=64, lanes_per_db=2, scratch_budget=budget) as sv:
        db, cap = sv.add_db(s0)
        assert cap == 3
        with pytest.raises(OnionError) as e:
            sv.add_db(s1)
        assert e.value.code == -3
        c, idx, q = make_queries(dbs, 0, 7, seed=901)
        tickets = [sv.submit(db, int(c[i]), q[i]) for i in range(7)]
        got = [sv.wait(t) for t in tickets]
        assert sv.stats(db)["mean_batch"] <= 3
    for i in range(7):
        assert np.array_equal(got[i], s0.answer(int(c[i]), q[i]))


def test_serve_one_by_one_database_and_discard():
    """A 1 x 1 database (no expansion levels, no further dimensions) in the
    serving loop; a discarded ticket is gone, the others still answer."""
    from onionpirv2_amd import GpuServer, OnionValueError
    from onionpirv2_amd.client import GpuClient
    from onionpirv2_amd.serve import GpuServe
    with GpuServer(1, 1, plain_bits=PLAIN_BITS) as srv:
        assert srv.info["expan_height"] == 0 and srv.info["poly_degree"] == N
        srv.db_generate(5, 0)
        c = GpuClient(1, 1, client_id=4, seed=44, plain_bits=PLAIN_BITS)
        srv.set_galois_key(4, c.galois_keys())
        srv.set_rgsw_key(4, c.rgsw_key())
        qs = [c.query(0, nonce=k) for k in range(3)]
        with GpuServe(max_batch=4, lanes_per_db=1, workers=1) as sv:
            db, cap = sv.add_db(srv)
            tickets = [sv.submit(db, 4, q) for q in qs]
            sv.discard(tickets[1])
            with pytest.raises(OnionValueError):
                sv.discard(tickets[1])
            for k in (0, 2):
                r = sv.wait(tickets[k])
                assert np.array_equal(r, srv.answer(4, qs[k]))
                assert np.array_equal(c.decrypt(r), record_of(5, 0, srv))
            with pytest.raises(OnionValueError):
                sv.wait(tickets[1], timeout_ms=0)


def record_of(seed, idx, srv):
    from onionpirv2_amd import dbgen
    return dbgen.plaintext(seed, 0, idx, t=srv.info["t"])


def test_serve_one_worker_serves_oldest_first(dbs):
    """One worker for two databases: the query that arrived first is answered
    in the first batch although the other database's queue fills behind it;
    every response equals the query answered alone."""
    from onionpirv2_amd.serve import GpuServe
    s0, s1 = dbs["servers"]
    c0, i0, q0 = make_queries(dbs, 0, 1, seed=960)
    c1, i1, q1 = make_queries(dbs, 1, 16, seed=961)
    with GpuServe(max_batch=2, lanes_per_db=1, workers=1) as sv:
        d0, _ = sv.add_db(s0)
        d1, _ = sv.add_db(s1)
        first = sv.submit(d0, int(c0[0]), q0[0])
        rest = [sv.submit(d1, int(c1[k]), q1[k]) for k in range(16)]
        r0 = sv.wait(first)
        assert sv.stats(d1)["completed"] <= 2  # at most the batch that ran beside nothing else
        got = [sv.wait(t, db_id=d1) for t in rest]
        assert sv.stats(d1)["batches"] >= 8
    assert np.array_equal(r0, s0.answer(int(c0[0]), q0[0]))
    for k in range(16):
        assert np.array_equal(got[k], s1.answer(int(c1[k]), q1[k]))


def test_serve_even_split_of_a_backlog(dbs):
    """batch_split = 1 on a database: a backlog above its cap (10 waiting at
    cap 4) is taken in equal shares; every response still equals the query
    answered alone."""
    from onionpirv2_amd.serve import GpuServe
    s0, s1 = dbs["servers"]
    saved = s0.get_tuning()
    c0, i0, q0 = make_queries(dbs, 0, 10, seed=970)
    c1, i1, q1 = make_queries(dbs, 1, 1, seed=971)
    try:
        s0.set_tuning(batch_split=1)
        with GpuServe(max_batch=4, lanes_per_db=1, workers=1) as sv:
            d0, cap = sv.add_db(s0)
            d1, _ = sv.add_db(s1)
            assert cap == 4
            busy = sv.submit(d1, int(c1[0]), q1[0])  # keeps the worker while the backlog forms
            ts = [sv.submit(d0, int(c0[k]), q0[k]) for k in range(10)]
            got = [sv.wait(t, db_id=d0) for t in ts]
            sv.wait(busy, db_id=d1)
            assert sv.stats(d0)["batches"] >= 3
    finally:
        s0.set_tuning(**saved)
    for k in range(10):
        assert np.array_equal(got[k], s0.answer(int(c0[k]), q0[k]))


def test_serve_poll_and_timeout(dbs):
    from onionpirv2_amd.serve import GpuServe, ServeTimeout
    srv = dbs["servers"][0]
    c, idx, q = make_queries(dbs, 0, 1, seed=902)
    with GpuServe(max_batch=4, lanes_per_db=1, linger_us=200000) as sv:
        sv.add_db(srv)
        t = sv.submit(0, int(c[0]), q[0])
        with pytest.raises(ServeTimeout):
            sv.wait(t, timeout_ms=0)  # the lane lingers 200 ms for more arrivals
        assert np.array_equal(sv.wait(t), srv.answer(int(c[0]), q[0]))


def test_shared_key_store_across_shapes(dbs):
    """Two new databases (same shapes and seeds as the fixture's) share one
    device key store, filled once with keys tall enough for both: answers
    equal those of the fixture's servers, which hold their own keys."""
    from onionpirv2_amd import GpuServer, OnionKeyError, OnionValueError
    from onionpirv2_amd.client import GpuClient
    (f0, o0), (f1, o1) = SHAPES
    a = GpuServer(f0, o0, plain_bits=PLAIN_BITS)   # h = 7
    b = GpuServer(f1, o1, plain_bits=PLAIN_BITS)   # h = 8
    a.db_generate(100, 0)
    b.db_generate(101, 0)
    b.share_keys(a)
    for cid in range(N_CLIENTS):
        tall = GpuClient(f1, o1, client_id=cid, seed=500 + cid, plain_bits=PLAIN_BITS)
        a.set_galois_key(cid, tall.galois_keys())   # 8 levels, uploaded through the h = 7 server
        a.set_rgsw_key(cid, tall.rgsw_key())
    for d, srv in enumerate((a, b)):
        cids, idx, qs = make_queries(dbs, d, 5, seed=950 + d)
        got = srv.answer_batch(cids, qs)
        ref = dbs["servers"][d]
        for i in range(5):
            assert np.array_equal(got[i], ref.answer(int(cids[i]), qs[i])), f"db {d} query {i}"
    # keys too short for the taller database
    short = GpuClient(f0, o0, client_id=77, seed=777, plain_bits=PLAIN_BITS)
    a.set_galois_key(77, short.galois_keys())
    a.set_rgsw_key(77, short.rgsw_key())
    q = GpuClient(f1, o1, client_id=77, seed=777, plain_bits=PLAIN_BITS).query(3, nonce=1)
    with pytest.raises(OnionKeyError):
        b.answer(77, q)
    a.answer(77, short.query(3, nonce=1))  # fine on the shorter database
    # a server with keys of its own cannot join a store
    c = GpuServer(f0, o0, plain_bits=PLAIN_BITS)
    c.set_galois_key(1, short.galois_keys())
    with pytest.raises(OnionValueError):
        c.share_keys(a)
    c.close()
    b.close()
    a.close()


def test_key_upload_while_batches_run_on_a_sharing_server(dbs):
    """Uploads through one server wait for batches in flight on another server
    of the same store; nothing deadlocks and every answer stays right."""
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    (f0, o0), (f1, o1) = SHAPES
    a = GpuServer(f0, o0, plain_bits=PLAIN_BITS)
    b = GpuServer(f1, o1, plain_bits=PLAIN_BITS)
    a.db_generate(100, 0)
    b.db_generate(101, 0)
    b.share_keys(a)
    tall = {cid: GpuClient(f1, o1, client_id=cid, seed=500 + cid, plain_bits=PLAIN_BITS) for cid in range(N_CLIENTS)}
    for cid, c in tall.items():
        a.set_galois_key(cid, c.galois_keys())
        a.set_rgsw_key(cid, c.rgsw_key())
    cids, idx, qs = make_queries(dbs, 1, 6, seed=990)
    ref = b.answer_batch(cids, qs)
    errors, results = [], []

    def batches():
        try:
            for _ in range(6):
                results.append(b.answer_batch(cids, qs))
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    def uploads():
        try:
            for k in range(6):
                c = tall[k % N_CLIENTS]
                a.set_galois_key(k % N_CLIENTS, c.galois_keys())
                a.set_rgsw_key(k % N_CLIENTS, c.rgsw_key())
                extra = GpuClient(f1, o1, client_id=200 + k, seed=2000 + k, plain_bits=PLAIN_BITS)
                a.set_galois_key(200 + k, extra.galois_keys())
                a.set_rgsw_key(200 + k, extra.rgsw_key())
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=batches), threading.Thread(target=uploads)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th), "deadlock"
    assert not errors, errors
    assert len(results) == 6 and all(np.array_equal(r, ref) for r in results)
    b.close()
    a.close()
