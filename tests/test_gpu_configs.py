"""The full-size database configurations under the GPU tests.

The reference's only correctness check is end-to-end record equality
(src/tests.cpp:111-134): the client decrypts exactly the entry it asked for.
Every full-size configuration gets that check here, plus responses equal to
the CPU oracle's bit for bit on the same keys, query and database:

  256x512    2 GiB NTT-form database: single answers
  512x1024   8 GiB: a batch of 64 queries from 64 clients in one pipeline,
             eager and replayed from a captured hipGraph
  128x256 + 256x1024 + 512x4096   36.5 GiB resident together: bursts of
             7 / 33 / 100 queries through the serving loop
  the same + 512x8192   100.5 GiB resident: 1000 clients' keys in one device
             key store, Poisson arrivals through the serving loop

Records are decrypted at t = 32771 (PlainMod 16, the reference's setting for
its N = 2048 presets); bit-exactness against the oracle is checked at
t = 32771 and at t = 65537 (PlainMod 17).  At t = 65537 the deeper shapes run
out of noise budget for some queries (a property of the parameters, the same
in the oracle), so there the oracle's response is the criterion.
"""
import concurrent.futures as cf
import os
import sys
import time

import numpy as np
import pytest

from oracle_lib import Oracle, ptr

pytestmark = pytest.mark.gpu

N = 2048


def say(msg):
    print(f"[configs] {msg}", file=sys.stderr, flush=True)


def oracle_db(o, seed):
    """The oracle's NTT-form database, built on up to 16 host threads."""
    db = np.zeros(o.num_pt * N, np.uint64)
    workers = max(1, min(16, os.cpu_count() or 1))
    chunk = (o.num_pt + workers - 1) // workers
    t0 = time.time()
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda w: o.L.orc_db_build_range(o.c, seed, 0, w * chunk, min(o.num_pt, (w + 1) * chunk),
                                                     ptr(db)), range(workers)))
    say(f"oracle database {o.fst}x{o.other} built in {time.time() - t0:.1f} s")
    return db


def oracle_answer(o, db, client, query):
    resp, _ = o.answer(db, client.galois_keys(), o.rgsw_key_ntt(client.rgsw_key()), query)
    return resp


def record(seed, idx, t):
    from onionpirv2_amd import dbgen
    return dbgen.plaintext(seed, 0, idx, t=t)


@pytest.mark.parametrize("plain_bits", [16, 17])
def test_cfg2_256x512_single_answers(plain_bits):
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    fst, other, seed = 256, 512, 202
    say(f"256x512 t-bits {plain_bits}")
    with GpuServer(fst, other, plain_bits=plain_bits) as srv:
        srv.db_generate(seed, 0)
        c = GpuClient(fst, other, client_id=3, seed=301, plain_bits=plain_bits)
        srv.set_galois_key(3, c.galois_keys())
        srv.set_rgsw_key(3, c.rgsw_key())
        num_pt, t = srv.info["num_pt"], srv.info["t"]
        idx = [0, num_pt - 1] + [int(i) for i in np.random.default_rng(plain_bits).integers(0, num_pt, 2)]
        qs = [c.query(i, nonce=10 + k) for k, i in enumerate(idx)]
        got = [srv.answer(3, q) for q in qs]
        o = Oracle(fst=fst, other=other, plain_bits=plain_bits)
        db = oracle_db(o, seed)
        for k in range(len(idx)):
            assert np.array_equal(got[k], oracle_answer(o, db, c, qs[k])), f"query {k} (index {idx[k]})"
        if plain_bits == 16:
            for k, i in enumerate(idx):
                assert np.array_equal(c.decrypt(got[k]), record(seed, i, t)), f"record {i}"


def _cfg3_setup(plain_bits, seed):
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    fst, other, B = 512, 1024, 64
    srv = GpuServer(fst, other, plain_bits=plain_bits)
    srv.db_generate(seed, 0)
    num_pt = srv.info["num_pt"]
    idx = np.random.default_rng(seed).choice(num_pt, size=B, replace=False)
    clients, qs = [], []
    for i in range(B):
        c = GpuClient(fst, other, client_id=i, seed=7000 + i, plain_bits=plain_bits)
        srv.set_galois_key(i, c.galois_keys())
        srv.set_rgsw_key(i, c.rgsw_key())
        clients.append(c)
        qs.append(c.query(int(idx[i]), nonce=i + 1))
    return srv, clients, idx, np.stack(qs), np.arange(B, dtype=np.uint64)


def test_cfg3_512x1024_batch_of_64():
    seed = 303
    say("512x1024 batch of 64, t = 32771")
    srv, clients, idx, qs, ids = _cfg3_setup(16, seed)
    try:
        t = srv.info["t"]
        resps = srv.answer_batch(ids, qs)
        bad = [i for i in range(64) if not np.array_equal(clients[i].decrypt(resps[i]), record(seed, int(idx[i]), t))]
        assert not bad, f"records not recovered for queries {bad}"
        # the same batch replayed from the captured pipeline
        srv.set_tuning(use_graph=1)
        for rep in range(2):
            assert np.array_equal(srv.answer_batch(ids, qs), resps), f"graph replay {rep}"
        srv.set_tuning(use_graph=0)
        o = Oracle(fst=512, other=1024, plain_bits=16)
        db = oracle_db(o, seed)
        for i in (0, 63):
            assert np.array_equal(resps[i], oracle_answer(o, db, clients[i], qs[i])), f"query {i}"
    finally:
        srv.close()


def test_cfg3_512x1024_batch_of_64_t65537():
    seed = 304
    say("512x1024 batch of 64, t = 65537")
    srv, clients, idx, qs, ids = _cfg3_setup(17, seed)
    try:
        resps = srv.answer_batch(ids, qs)
        o = Oracle(fst=512, other=1024, plain_bits=17)
        db = oracle_db(o, seed)
        for i in (0, 63):
            assert np.array_equal(resps[i], oracle_answer(o, db, clients[i], qs[i])), f"query {i}"
        assert np.array_equal(srv.answer(5, qs[5]), resps[5])
    finally:
        srv.close()


CFG4_SHAPES = [(128, 256), (256, 1024), (512, 4096)]


def test_cfg4_three_databases_bursts():
    """Bursts of 7 / 33 / 100 queries spread over three resident databases go
    through the serving loop; each response equals the same query answered
    alone, decrypts to its record, and one per database equals the oracle's."""
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    from onionpirv2_amd.serve import GpuServe
    say("three databases, bursts of 7 / 33 / 100")
    servers, clients = [], {}
    try:
        for d, (fst, other) in enumerate(CFG4_SHAPES):
            s = GpuServer(fst, other, plain_bits=16)
            s.db_generate(400 + d, 0)
            for cid in range(6):
                c = GpuClient(fst, other, client_id=cid, seed=4000 + 10 * d + cid, plain_bits=16)
                s.set_galois_key(cid, c.galois_keys())
                s.set_rgsw_key(cid, c.rgsw_key())
                clients[(d, cid)] = c
            servers.append(s)
        rng = np.random.default_rng(44)
        anchors = {}
        with GpuServe(max_batch=64, lanes_per_db=1) as sv:
            for s in servers:
                sv.add_db(s)
            for burst in (7, 33, 100):
                work = []
                for k in range(burst):
                    d = k % 3
                    cid = int(rng.integers(6))
                    i = int(rng.integers(servers[d].info["num_pt"]))
                    work.append((d, cid, i, clients[(d, cid)].query(i, nonce=burst * 1000 + k)))
                tickets = [sv.submit(d, cid, q) for d, cid, _, q in work]
                resps = [sv.wait(tk, db_id=w[0]) for tk, w in zip(tickets, work)]
                for (d, cid, i, q), r in zip(work, resps):
                    assert np.array_equal(r, servers[d].answer(cid, q)), f"burst {burst}: db {d} record {i}"
                    assert np.array_equal(clients[(d, cid)].decrypt(r), record(400 + d, i, servers[d].info["t"]))
                    anchors.setdefault(d, (cid, q, r))
        for d, (fst, other) in enumerate(CFG4_SHAPES):
            o = Oracle(fst=fst, other=other, plain_bits=16)
            db = oracle_db(o, 400 + d)
            cid, q, r = anchors[d]
            assert np.array_equal(r, oracle_answer(o, db, clients[(d, cid)], q)), f"db {fst}x{other}"
            del db
    finally:
        for s in servers:
            s.close()


def test_cfg5_four_databases_1000_clients():
    """Four databases resident together (100.5 GiB) sharing one device key
    store with 1000 clients' keys; Poisson arrivals through the serving loop
    to a random database, client and record; every response decrypts to its
    record; one 512x8192 answer equals the oracle's."""
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    from onionpirv2_amd.serve import GpuServe
    shapes = CFG4_SHAPES + [(512, 8192)]
    n_clients = 1000
    say("four databases (100.5 GiB), 1000 clients")
    servers = []
    try:
        for d, (fst, other) in enumerate(shapes):
            s = GpuServer(fst, other, plain_bits=16)
            s.db_generate(500 + d, 0)
            if servers:
                s.share_keys(servers[0])
            servers.append(s)
        assert sum(s.info["num_pt"] for s in servers) * N * 8 == int(100.5 * 2**30)
        tall = shapes[-1]  # keys for the tallest expansion serve every database
        t0 = time.time()
        for cid in range(n_clients):
            c = GpuClient(tall[0], tall[1], client_id=cid, seed=50000 + cid, plain_bits=16)
            servers[0].set_galois_key(cid, c.galois_keys())
            servers[0].set_rgsw_key(cid, c.rgsw_key())
        say(f"1000 clients keyed in {time.time() - t0:.1f} s")
        rng = np.random.default_rng(55)
        qclients = {}

        def client(d, cid):
            if (d, cid) not in qclients:
                qclients[(d, cid)] = GpuClient(shapes[d][0], shapes[d][1], client_id=cid, seed=50000 + cid,
                                               plain_bits=16)
            return qclients[(d, cid)]

        n, rate = 160, 400.0
        work = []
        for k in range(n):
            d = int(rng.integers(4))
            cid = int(rng.integers(n_clients))
            i = int(rng.integers(servers[d].info["num_pt"]))
            work.append((d, cid, i, client(d, cid).query(i, nonce=k + 1)))
        gaps = rng.exponential(1.0 / rate, size=n)
        with GpuServe(max_batch=64, lanes_per_db=1, scratch_budget=120 << 30) as sv:
            for s in servers:
                sv.add_db(s)
            tickets = []
            due = time.perf_counter()
            for k, (d, cid, i, q) in enumerate(work):
                due += gaps[k]
                while time.perf_counter() < due:
                    pass
                tickets.append(sv.submit(d, cid, q))
            for (d, cid, i, q), tk in zip(work, tickets):
                r = sv.wait(tk, db_id=d)
                assert np.array_equal(client(d, cid).decrypt(r), record(500 + d, i, 32771)), \
                    f"db {shapes[d]} client {cid} record {i}"
            st = sv.stats()
            assert st["completed"] == n and st["failed"] == 0
        # one answer on the largest database against the oracle
        d = 3
        cid, i = 123, int(rng.integers(servers[d].info["num_pt"]))
        q = client(d, cid).query(i, nonce=999)
        got = servers[d].answer(cid, q)
        o = Oracle(fst=512, other=8192, plain_bits=16)
        db = oracle_db(o, 500 + d)
        assert np.array_equal(got, oracle_answer(o, db, client(d, cid), q))
    finally:
        for s in reversed(servers):
            s.close()
