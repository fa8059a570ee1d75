"""The device path reproduces the committed golden fixtures (tests/golden/)
stage by stage: expansion, conversion, first dimension, every multiplexer
level, the response and its packed form, through the stage taps and the full
answer."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CASES = json.load(open(os.path.join(GOLDEN, "golden.json")))["cases"]
N = 2048


def sha(a) -> str:
    if isinstance(a, (bytes, bytearray)):
        return hashlib.sha256(a).hexdigest()
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_device_reproduces_golden_stages(case):
    from onionpirv2_amd import GpuServer
    h = case["sha256"]
    o = Oracle(fst=case["fst"], other=case["other"], plain_bits=case["plain_bits"])
    sk = o.keygen(case["client_seed"])
    gk, rk = o.galois_keys(case["client_seed"], sk), o.rgsw_key(case["client_seed"], sk)
    assert sha(gk) == h["galois_keys"] and sha(rk) == h["rgsw_key"]
    q = o.query(case["client_seed"], case["nonce"], sk, case["index"])
    assert sha(q) == h["query"]
    with GpuServer(case["fst"], case["other"], plain_bits=case["plain_bits"]) as srv:
        srv.db_generate(case["db_seed"], case["db_mode"])
        srv.set_galois_key(1, gk)
        srv.set_rgsw_key(1, rk)
        leaves = srv.tap_expand(1, q)
        assert sha(leaves) == h["leaves"], "expansion"
        rgsw = srv.tap_convert(1, leaves)
        assert sha(rgsw) == h["rgsw"], "conversion"
        fd = srv.tap_first_dim(leaves[: case["fst"] * 2 * N])
        assert sha(fd) == h["first_dim"], "first dimension"
        cts, per_level = fd, 2 * o.l * 2 * N
        for d in range(o.ndim - 1):
            cts = srv.tap_mux(cts, rgsw[d * per_level:(d + 1) * per_level])
            assert sha(cts) == h[f"mux_{d}"], f"mux level {d}"
        resp = srv.answer(1, q)
        assert sha(resp) == h["response"], "response"
        assert sha(srv.pack_response(resp)) == h["packed"], "packed response"
        # the batched path (matrix-core first dimension) gives the same bits
        got = srv.answer_batch([1, 1, 1], np.stack([q, q, q]))
        assert all(sha(g) == h["response"] for g in got)
    if case["name"] == "cfg1_256x32_t65537":
        assert np.array_equal(resp, np.load(os.path.join(GOLDEN, f"{case['name']}_response.npy")))
