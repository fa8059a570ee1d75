"""The product client (libonionpir.so, CPU) against the oracle's client: same
seeds give bit-identical keys and queries, and both decrypt the oracle
server's response to the stored record."""
import numpy as np
import pytest

from oracle_lib import Oracle

FST, OTHER, SEED = 256, 32, 11


@pytest.fixture(scope="module")
def pair():
    from onionpirv2_amd.client import GpuClient
    o = Oracle(fst=FST, other=OTHER)
    c = GpuClient(FST, OTHER, client_id=3, seed=SEED)
    return o, c, o.keygen(SEED)


def test_galois_keys_bit_identical(pair):
    o, c, sk = pair
    assert np.array_equal(c.galois_keys(), o.galois_keys(SEED, sk))


def test_rgsw_key_bit_identical(pair):
    o, c, sk = pair
    assert np.array_equal(c.rgsw_key(), o.rgsw_key(SEED, sk))


@pytest.mark.parametrize("idx", [0, 77, 19 * 256 + 77, FST * OTHER - 1])
def test_query_bit_identical(pair, idx):
    o, c, sk = pair
    assert np.array_equal(c.query(idx, nonce=5 + idx), o.query(SEED, 5 + idx, sk, idx))


def test_query_index_out_of_range(pair):
    from onionpirv2_amd import OnionValueError
    _, c, _ = pair
    with pytest.raises(OnionValueError):
        c.query(FST * OTHER)


def test_decrypt_and_entry_against_oracle_server(pair):
    o, c, sk = pair
    db = o.db_build(7, 1)
    rkn = o.rgsw_key_ntt(c.rgsw_key())
    gk = c.galois_keys()
    for idx in (5, 4000):
        resp, _ = o.answer(db, gk, rkn, c.query(idx, nonce=idx))
        wire = o.pack(resp)
        r2 = c.unpack_response(wire)
        assert np.array_equal(r2, o.unpack(wire))
        pt = c.decrypt(r2)
        assert np.array_equal(pt, o.db_plaintext(7, 1, idx))
        e = c.entry(idx, pt)
        assert e == o.entry(pt) and e[:8] == idx.to_bytes(8, "big")


def test_truncated_response_raises(pair):
    from onionpirv2_amd import OnionValueError
    _, c, _ = pair
    with pytest.raises(OnionValueError):
        c.unpack_response(b"\x00" * 100)


def test_keys_of_a_smaller_shape_are_a_prefix():
    """A client's Galois key of level a (element 1 + (N >> a)) and its RGSW key do
    not depend on the database shape, so keys generated for the tallest
    expansion serve every database that shares a device key store."""
    from onionpirv2_amd.client import GpuClient
    small = GpuClient(128, 256, client_id=3, seed=42, plain_bits=16)   # h = 8
    big = GpuClient(512, 8192, client_id=3, seed=42, plain_bits=16)    # h = 10
    gs, gb = small.galois_keys(), big.galois_keys()
    assert gs.size == 8 * 4 * 2048 and gb.size == 10 * 4 * 2048
    assert np.array_equal(gb[: gs.size], gs)
    assert np.array_equal(small.rgsw_key(), big.rgsw_key())
