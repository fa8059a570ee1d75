"""onion_gpu_answer_batch_noalloc: the batch answer with its scratch in the
library's static device array instead of lane buffers.  Its responses must
equal onion_gpu_answer_batch's bit for bit, call after call, whatever the
previous call left in the array; batches that do not fit the array are
refused."""
import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N = 2048


def keyed_server(fst, other, seed, n_clients):
    from onionpirv2_amd import GpuServer
    from onionpirv2_amd.client import GpuClient
    srv = GpuServer(fst, other, plain_bits=16)
    srv.db_generate(seed, 0)
    clients = []
    for cid in range(n_clients):
        c = GpuClient(fst, other, client_id=cid, seed=seed * 100 + cid, plain_bits=16)
        srv.set_galois_key(cid, c.galois_keys())
        srv.set_rgsw_key(cid, c.rgsw_key())
        clients.append(c)
    return srv, clients


def test_noalloc_matches_answer_batch_on_successive_calls():
    from onionpirv2_amd import dbgen
    srv, clients = keyed_server(128, 64, 61, 3)
    try:
        rng = np.random.default_rng(6)
        idx = rng.integers(0, srv.info["num_pt"], size=5)
        qs = np.stack([clients[k % 3].query(int(i), nonce=100 + k) for k, i in enumerate(idx)])
        ids = np.array([k % 3 for k in range(5)], dtype=np.uint64)
        # first call: three queries; second call: one query (a smaller input)
        for sel in (slice(0, 3), slice(3, 4), slice(4, 5)):
            got = srv.answer_batch_noalloc(ids[sel], qs[sel])
            want = srv.answer_batch(ids[sel], qs[sel])
            assert np.array_equal(got, want), f"queries {sel}"
            for k in range(sel.start, sel.stop):
                pt = clients[k % 3].decrypt(got[k - sel.start])
                assert np.array_equal(pt, dbgen.plaintext(61, 0, int(idx[k]), t=srv.info["t"]))
        # anchored to the oracle once
        o = Oracle(fst=128, other=64, plain_bits=16)
        c = clients[1]
        exp, _ = o.answer(o.db_build(61, 0), c.galois_keys(), o.rgsw_key_ntt(c.rgsw_key()), qs[1])
        assert np.array_equal(srv.answer_batch_noalloc(ids[1:2], qs[1:2])[0], exp)
    finally:
        srv.close()


def test_noalloc_across_databases_and_limit():
    """A 512x1024 query (the array's size was chosen for one of them), then a
    query on a smaller database; two 512x1024 queries do not fit."""
    from onionpirv2_amd import OnionValueError
    big, bc = keyed_server(512, 1024, 62, 2)
    small, sc = keyed_server(64, 16, 63, 1)
    try:
        qb = np.stack([bc[0].query(12345, nonce=1), bc[1].query(7, nonce=1)])
        ids = np.array([0, 1], dtype=np.uint64)
        assert np.array_equal(big.answer_batch_noalloc(ids[:1], qb[:1]), big.answer_batch(ids[:1], qb[:1]))
        qs = sc[0].query(99, nonce=1)[None, :]
        z = np.zeros(1, np.uint64)
        assert np.array_equal(small.answer_batch_noalloc(z, qs), small.answer_batch(z, qs))
        with pytest.raises(OnionValueError, match="static scratch"):
            big.answer_batch_noalloc(ids, qb)
        assert np.array_equal(big.answer_batch_noalloc(ids[1:], qb[1:]), big.answer_batch(ids[1:], qb[1:]))
    finally:
        small.close()
        big.close()
