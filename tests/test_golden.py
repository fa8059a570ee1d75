"""The oracle reproduces the committed golden fixtures (tests/golden/): every
stage's SHA-256 for both cases, and the full query, response and packed
response of the first.  A change to the oracle's arithmetic fails here."""
import importlib.util
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _make_golden():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


CASES = json.load(open(os.path.join(GOLDEN, "golden.json")))["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_golden_hashes(case):
    mg = _make_golden()
    st = mg.stages({k: v for k, v in case.items() if k != "sha256"})
    got = {k: mg.sha(v) for k, v in st.items()}
    assert got == case["sha256"]


def test_full_vectors_and_product_client():
    from onionpirv2_amd.client import GpuClient
    case = next(c for c in CASES if c["name"] == "cfg1_256x32_t65537")
    q = np.load(os.path.join(GOLDEN, f"{case['name']}_query.npy"))
    resp = np.load(os.path.join(GOLDEN, f"{case['name']}_response.npy"))
    packed = open(os.path.join(GOLDEN, f"{case['name']}_response.bin"), "rb").read()
    mg = _make_golden()
    assert mg.sha(q) == case["sha256"]["query"]
    assert mg.sha(resp) == case["sha256"]["response"]
    assert mg.sha(packed) == case["sha256"]["packed"] and len(packed) == 13824
    # the product client builds the same query and unpacks the same response
    c = GpuClient(case["fst"], case["other"], client_id=1, seed=case["client_seed"], plain_bits=case["plain_bits"])
    assert np.array_equal(c.query(case["index"], nonce=case["nonce"]), q)
    assert np.array_equal(c.unpack_response(packed), resp)
