"""Write the golden fixtures of this directory from the CPU oracle.

The reference ships no golden vectors or known-answer tests.  These are this
project's own: for each case, the seeds, the SHA-256 of every stage's output
(database, keys, query, expansion leaves, converted RGSWs, first-dimension
output, each multiplexer level, response, packed response, decrypted record)
and, for the first case, the query and response in full.  They pin drift of
the oracle (test_golden.py) and of the device path (test_gpu_golden.py)
against a fixed record; fidelity to SEAL is pinned elsewhere (test_oracle.py).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle  # noqa: E402

N = 2048

CASES = [
    dict(name="cfg1_256x32_t65537", fst=256, other=32, plain_bits=17, db_seed=7, db_mode=0,
         client_seed=11, nonce=123, index=4242),
    dict(name="entries_128x16_t32771", fst=128, other=16, plain_bits=16, db_seed=21, db_mode=1,
         client_seed=31, nonce=7, index=1000),
]
FULL_VECTORS = "cfg1_256x32_t65537"


def sha(a) -> str:
    if isinstance(a, (bytes, bytearray)):
        return hashlib.sha256(a).hexdigest()
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def stages(case: dict) -> dict:
    """Every stage of one answer through the oracle, as arrays and bytes."""
    o = Oracle(fst=case["fst"], other=case["other"], plain_bits=case["plain_bits"])
    db = o.db_build(case["db_seed"], case["db_mode"])
    sk = o.keygen(case["client_seed"])
    gk = o.galois_keys(case["client_seed"], sk)
    rk = o.rgsw_key(case["client_seed"], sk)
    rkn = o.rgsw_key_ntt(rk)
    q = o.query(case["client_seed"], case["nonce"], sk, case["index"])
    leaves = o.expand(q, gk)
    rgsw = o.convert(leaves, rkn)
    fd = o.first_dim(db, leaves[: o.fst * 2 * N])
    out = dict(db=db, galois_keys=gk, rgsw_key=rk, query=q, leaves=leaves, rgsw=rgsw, first_dim=fd)
    cts = fd.reshape(o.other, 2 * N)
    per_level = 2 * o.l * 2 * N
    for d in range(o.ndim - 1):
        cts = o.mux_level(cts.reshape(-1), rgsw[d * per_level:(d + 1) * per_level]).reshape(-1, 2 * N)
        out[f"mux_{d}"] = cts.reshape(-1)
    resp, _ = o.answer(db, gk, rkn, q)
    switched = np.array([o.L.orc_mod_switch_one(int(v), o.q, o.small_q) for v in cts.reshape(-1)],
                        dtype=np.uint64)
    assert np.array_equal(switched, resp), "stage chain disagrees with the oracle's answer"
    out["response"] = resp
    out["packed"] = o.pack(resp)
    pt = o.decrypt_small(sk, resp)
    assert np.array_equal(pt, o.db_plaintext(case["db_seed"], case["db_mode"], case["index"])), "record not recovered"
    out["plaintext"] = pt
    return out


def main() -> None:
    golden = {"note": "SHA-256 of each stage's output (uint64 arrays little-endian, packed response as bytes)",
              "cases": []}
    for case in CASES:
        st = stages(case)
        golden["cases"].append(dict(case, sha256={k: sha(v) for k, v in st.items()}))
        if case["name"] == FULL_VECTORS:
            np.save(os.path.join(HERE, f"{case['name']}_query.npy"), st["query"])
            np.save(os.path.join(HERE, f"{case['name']}_response.npy"), st["response"])
            with open(os.path.join(HERE, f"{case['name']}_response.bin"), "wb") as f:
                f.write(st["packed"])
        print(case["name"], "ok", flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
