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
        resp, _ = o.answer(db, gk, rkn, c.query(idx, nonce=10000 + idx))
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


def test_chacha20_block_known_answer():
    """RFC 8439 section 2.3.2: key 00..1f, block counter 1, nonce
    00:00:00:09:00:00:00:4a:00:00:00:00 (state words 12..15 = 1, 0x09000000,
    0x4a000000, 0)."""
    from onionpirv2_amd.client import chacha20_block
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    got = chacha20_block(key, [1, 0x09000000, 0x4a000000, 0])
    want = bytes.fromhex(
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert b"".join(w.to_bytes(4, "little") for w in got) == want


def test_secure_clients_differ_and_decrypt():
    """Two clients without a seed draw independent keys (ChaCha20 keyed from
    the operating system); each still decrypts the oracle server's answer."""
    from onionpirv2_amd.client import GpuClient
    o = Oracle(fst=16, other=4)
    a, b = GpuClient(16, 4, client_id=1), GpuClient(16, 4, client_id=2)
    ga, gb = a.galois_keys(), b.galois_keys()
    assert not np.array_equal(ga, gb)
    db = o.db_build(9, 0)
    for c, gk in ((a, ga), (b, gb)):
        resp, _ = o.answer(db, gk, o.rgsw_key_ntt(c.rgsw_key()), c.query(21))
        assert np.array_equal(c.decrypt(resp), o.db_plaintext(9, 0, 21))


def test_query_nonces_are_never_reused():
    from onionpirv2_amd import OnionValueError
    from onionpirv2_amd.client import GpuClient
    c = GpuClient(16, 4, client_id=1, seed=5)
    q1, q2 = c.query(3), c.query(3)  # default nonces: 0, then 1
    assert not np.array_equal(q1[2048:], q2[2048:])
    with pytest.raises(OnionValueError, match="already used"):
        c.query(3, nonce=1)
    with pytest.raises(OnionValueError, match="2\\^48"):
        c.query(3, nonce=1 << 48)
    c.query(3, nonce=(1 << 48) - 1)
    q3 = c.query(3)  # skips the used nonces 0 and 1
    assert not any(np.array_equal(q3[2048:], q[2048:]) for q in (q1, q2))


def test_query_streams_never_reach_key_streams():
    """No query nonce reproduces a key's uniform polynomial: the query's c1 is
    compared with the c1 of every Galois key (limb q, back in coefficient form)
    and every RGSW key row, for the nonces next to the key rows' own numbers
    (0x10000 + level, 0x20000 + row), the first 256 and the largest allowed."""
    from onionpirv2_amd.client import GpuClient
    o = Oracle(fst=FST, other=OTHER)
    c = GpuClient(FST, OTHER, client_id=3, seed=SEED)
    n = o.N
    gk = c.galois_keys().reshape(o.h, 2, 2, n)
    rk = c.rgsw_key().reshape(2 * o.l_key, 2, n)
    key_c1 = {tuple(o.intt(gk[a, 1, 0])) for a in range(o.h)} | {tuple(rk[r, 1]) for r in range(2 * o.l_key)}
    assert len(key_c1) == o.h + 2 * o.l_key
    nonces = list(range(256)) + [0x10000 + a for a in range(16)] + [0x20000 + r for r in range(40)] + \
        [(1 << 17) - 1, (1 << 48) - 1]
    for nonce in nonces:
        q = c.query(nonce % o.num_pt, nonce=nonce)
        assert tuple(q[n:]) not in key_c1, nonce
