"""Pins the CPU oracle (oracle/*.c) before it is trusted as the checker.

The reference ships no golden vectors or known-answer tests (its only check is
end-to-end record retrieval, src/tests.cpp:111-134), so each oracle function
is pinned here by what the mathematics fixes: independent recomputation in
Python integers (brute force on small inputs), closed forms, and invariants.
Each check is chosen so that a plausible slip (a dropped term, a wrong sign
or index, a transposed operand) fails it.
"""
import random
from fractions import Fraction

import numpy as np
import pytest
from sympy import isprime

from oracle_lib import Oracle, ptr, u64

Q = 1152921504606748673
P = 1152921504606830593
SMALL_Q = 134176769


def bitrev(x, n):
    return int(format(x, f"0{n}b")[::-1], 2) if n else 0


def negacyclic_mul(a, b, m):
    n = len(a)
    out = [0] * n
    for i, ai in enumerate(a):
        if ai == 0:
            continue
        for j, bj in enumerate(b):
            k = i + j
            if k < n:
                out[k] = (out[k] + ai * bj) % m
            else:
                out[k - n] = (out[k - n] - ai * bj) % m
    return out


def min_root_python(m, n):
    """Smallest primitive 2n-th root of unity mod m, by enumeration of all of them."""
    for x in range(2, 1000):
        r = pow(x, (m - 1) // (2 * n), m)
        if pow(r, n, m) == m - 1:
            return min(pow(r, k, m) for k in range(1, 2 * n, 2))
    raise AssertionError


# ---------------------------------------------------------------------------------
# parameters (SEAL CoeffModulus::Create rules, utils::generate_prime, PirParams)
# ---------------------------------------------------------------------------------

def test_moduli_follow_seal_rules():
    o = Oracle()
    assert (o.q, o.p, o.small_q, o.t) == (Q, P, SMALL_Q, 65537)
    for m, bits in ((Q, 60), (P, 60), (SMALL_Q, 27)):
        assert isprime(m) and m % 4096 == 1 and m.bit_length() == bits
    # p is the largest 60-bit prime = 1 mod 4096, q the next one down
    for v in range(2**60 - 4095, P, -4096):
        assert not isprime(v)
    for v in range(P - 4096, Q, -4096):
        assert not isprime(v)
    for v in range(2**27 - 4095, SMALL_Q, -4096):
        assert not isprime(v)
    # generate_prime: smallest prime above 2^(bits-1); PlainMod 16 -> 32771
    assert o.L.orc_generate_prime(16) == 32771
    assert o.L.orc_generate_prime(17) == 65537
    # gadget bases: ceil(60 / l)
    assert (o.base_log2, o.base_log2_key) == (12, 6)
    assert (o.ndim, o.u, o.h) == (6, 281, 9)


def test_shape_rule_matches_reference():
    o = Oracle()
    fst, other = u64(1), u64(1)
    # NumEntries 2^16, 1 entry per plaintext, MaxFstDimSz 256 (database_constants.h:215-223)
    o.L.orc_shape_from_entries(1 << 16, 1, 256, ptr(fst), ptr(other))
    assert (int(fst[0]), int(other[0])) == (256, 256)
    o.L.orc_shape_from_entries(1000, 1, 256, ptr(fst), ptr(other))
    assert (int(fst[0]), int(other[0])) == (250, 4)


def test_barrett128_matches_integer_remainder():
    o = Oracle()
    rng = random.Random(1)
    vals = [0, 1, Q - 1, Q, Q + 1, 2**64 - 1, 2**64, 2**128 - 1, (Q - 1) ** 2,
            512 * (Q - 1) ** 2 % 2**128, Q * (2**67 + 3)]
    vals += [rng.getrandbits(128) for _ in range(2000)] + [rng.getrandbits(121) for _ in range(500)]
    for which, m in ((0, Q), (1, P), (2, SMALL_Q)):
        for v in vals:
            assert o.barrett(which, v) == v % m, (which, v)


# ---------------------------------------------------------------------------------
# NTT (SEAL ntt_negacyclic_harvey order, minimal primitive root)
# ---------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [16, 32])
def test_ntt_matches_definition_small(n):
    o = Oracle(N=n, fst=2, other=1, l=1, l_key=1)
    for which, m in ((0, o.q), (1, o.p), (2, o.small_q)):
        psi = min_root_python(m, n)
        assert o.L.orc_ntt_psi(o.c, which) == psi
        logn = n.bit_length() - 1
        rng = random.Random(n + which)
        a = [rng.randrange(m) for _ in range(n)]
        exp = [sum(a[i] * pow(psi, (2 * bitrev(j, logn) + 1) * i, m) for i in range(n)) % m
               for j in range(n)]
        assert list(map(int, o.ntt(a, which))) == exp
        assert list(map(int, o.intt(exp, which))) == a


def test_ntt_2048_matches_naive_and_inverts():
    o = Oracle()
    rng = np.random.default_rng(2)
    for which, m in ((0, o.q), (1, o.p), (2, o.small_q)):
        a = rng.integers(0, m, size=2048, dtype=np.uint64)
        f = o.ntt(a, which)
        assert np.array_equal(f, o.ntt_naive(a, which))
        assert np.array_equal(o.intt(f, which), a)
        assert int(f.max()) < m


def test_ntt_convolution_theorem():
    o = Oracle(N=64, fst=4, other=2, l=1, l_key=1)
    rng = random.Random(5)
    a = [rng.randrange(o.q) for _ in range(64)]
    b = [rng.randrange(o.q) for _ in range(64)]
    fa, fb = o.ntt(a), o.ntt(b)
    prod = np.array([int(x) * int(y) % o.q for x, y in zip(fa, fb)], dtype=np.uint64)
    assert list(map(int, o.intt(prod))) == negacyclic_mul(a, b, o.q)


# ---------------------------------------------------------------------------------
# Galois automorphism, x^-k shift
# ---------------------------------------------------------------------------------

def galois(o, a, g, m):
    out = u64(o.N)
    x = np.ascontiguousarray(a, dtype=np.uint64)
    o.L.orc_galois_coeff(o.c, ptr(x), g, m, ptr(out))
    return out


def test_galois_is_substitution_and_composes():
    o = Oracle(N=32, fst=4, other=2, l=1, l_key=1)
    rng = random.Random(3)
    a = [rng.randrange(o.q) for _ in range(32)]
    a[5] = 0
    for g in (3, 5, 17, 33, 63):
        exp = [0] * 32  # brute force a(x^g) mod (x^32 + 1)
        for i, ai in enumerate(a):
            e = (i * g) % 64
            if e < 32:
                exp[e] = (exp[e] + ai) % o.q
            else:
                exp[e - 32] = (exp[e - 32] - ai) % o.q
        assert list(map(int, galois(o, a, g, o.q))) == exp
    for g1, g2 in ((3, 5), (17, 33)):
        assert np.array_equal(galois(o, galois(o, a, g1, o.q), g2, o.q), galois(o, a, g1 * g2 % 64, o.q))


def test_shift_is_multiplication_by_x_minus_k():
    o = Oracle(N=32, fst=4, other=2, l=1, l_key=1)
    rng = random.Random(4)
    a = [rng.randrange(o.q) for _ in range(32)]
    a[0] = 0
    for k in (1, 2, 8, 16):
        out = u64(32)
        o.L.orc_shift_neg(o.c, ptr(np.array(a, dtype=np.uint64)), k, ptr(out))
        xk = [0] * 32
        xk[32 - k] = o.q - 1  # x^-k = -x^(N-k)
        assert list(map(int, out)) == negacyclic_mul(a, xk, o.q)


def test_ntt_form_galois_permutation_and_shift_factor():
    """The slot identities the NTT-form expansion levels use, on the oracle's
    2048-point transform (slot s holds a(psi^o(s)), o(s) = 2 bitrev(s) + 1):
    NTT(sigma_g a)[s] = NTT(a)[bitrev((g o(s) mod 2N) >> 1)] and
    NTT(x^-k a)[s] = psi^(-k o(s)) NTT(a)[s], at every level's g = N/k + 1."""
    n = 2048
    o = Oracle(N=n, fst=64, other=2)
    q = o.q
    psi = int(o.L.orc_ntt_psi(o.c, 0))
    rng = np.random.default_rng(12)
    a = rng.integers(0, q, n, dtype=np.uint64)
    A = o.ntt(a)
    br = np.array([int(format(i, "011b")[::-1], 2) for i in range(n)])
    odd = 2 * br + 1
    for level in range(11):
        k = 1 << level
        g = n // k + 1
        src = br[((g * odd) % (2 * n)) >> 1]
        assert np.array_equal(o.ntt(galois(o, a, g, q)), A[src]), f"sigma, level {level}"
        shifted = u64(n)
        o.L.orc_shift_neg(o.c, ptr(a), k, ptr(shifted))
        base = pow(pow(psi, k, q), q - 2, q)
        exp = [int(A[s]) * pow(base, int(odd[s]), q) % q for s in range(n)]
        assert list(map(int, o.ntt(shifted))) == exp, f"x^-k, level {level}"


# ---------------------------------------------------------------------------------
# key switching with the special prime
# ---------------------------------------------------------------------------------

def test_keyswitch_is_rounded_division_by_p():
    """ct_k += floor((X + floor(p/2)) / p) mod q, X = CRT of target*key over qp."""
    n = 32
    o = Oracle(N=n, fst=4, other=2, l=1, l_key=1)
    q, p = o.q, o.p
    rng = random.Random(6)
    target = [rng.randrange(q) for _ in range(n)]
    key_coeff = [[rng.randrange(q * p) for _ in range(n)] for _ in range(2)]
    key = np.zeros(4 * n, dtype=np.uint64)
    for comp in range(2):
        key[(2 * comp) * n:(2 * comp + 1) * n] = o.ntt([v % q for v in key_coeff[comp]], 0)
        key[(2 * comp + 1) * n:(2 * comp + 2) * n] = o.ntt([v % p for v in key_coeff[comp]], 1)
    ct = np.array([rng.randrange(q) for _ in range(2 * n)], dtype=np.uint64)
    got = ct.copy()
    o.L.orc_keyswitch_add(o.c, ptr(np.array(target, dtype=np.uint64)), ptr(key), ptr(got))
    for comp in range(2):
        X = negacyclic_mul(target, key_coeff[comp], q * p)
        for i in range(n):
            add = ((X[i] + p // 2) // p) % q
            assert int(got[comp * n + i]) == (int(ct[comp * n + i]) + add) % q


def test_apply_galois_decrypts_to_permuted_message():
    o = Oracle()
    sk = o.keygen(31)
    gk = o.galois_keys(31, sk)
    msg = np.array([(i * 37) % o.t for i in range(o.N)], dtype=np.uint64)
    ct = o.encrypt(31, 5, sk, msg)
    for level in (0, 3, o.h - 1):
        g = o.N // (1 << level) + 1
        out = u64(2 * o.N)
        o.L.orc_apply_galois(o.c, ptr(ct), g, ptr(gk[level * 4 * o.N:(level + 1) * 4 * o.N].copy()), ptr(out))
        assert np.array_equal(o.decrypt(sk, out), galois(o, msg, g, o.t))


# ---------------------------------------------------------------------------------
# decomposition and external product
# ---------------------------------------------------------------------------------

def test_decomposition_recomposes_exactly():
    o = Oracle()
    rng = np.random.default_rng(8)
    ct = rng.integers(0, o.q, size=2 * o.N, dtype=np.uint64)
    ct[:3] = [0, o.q - 1, 2**59]
    for l, B in ((o.l, o.base_log2), (o.l_key, o.base_log2_key)):
        d = u64(2 * l * o.N)
        o.L.orc_decompose(o.c, ptr(ct), l, B, ptr(d))
        d = d.reshape(2 * l, o.N)
        assert int(d.max()) < 2**B
        gad = [2 ** ((l - 1 - j) * B) for j in range(l)]
        for poly in range(2):
            for i in (0, 1, 2, 100, 2047):
                assert sum(int(d[poly * l + j, i]) * gad[j] for j in range(l)) == int(ct[poly * o.N + i])


def test_external_product_with_folded_last_digit():
    """The identity the device's NTT-form multiplexer relies on: with each
    polynomial's RGSW rows j < l-1 replaced by G_j - 2^((l-1-j)B) G_{l-1}
    (mod q), the last digit's transform is never needed,
      sum_j NTT(d_j) G_j = sum_{j<l-1} NTT(d_j) G'_j + NTT(c) G_{l-1}  (mod q),
    because c = sum_j d_j 2^((l-1-j)B) exactly.  Checked against the oracle's
    external product on a random ciphertext, in Python integers."""
    o = Oracle()
    q, N, l, B = o.q, o.N, o.l, o.base_log2
    sk = o.keygen(43)
    g = o.rgsw(43, 5, sk, 1).reshape(2 * l, 2, N).astype(object)
    rng = np.random.default_rng(12)
    ct = rng.integers(0, q, size=2 * N, dtype=np.uint64)
    ct[:2] = [q - 1, 0]
    exp = u64(2 * N)
    o.L.orc_external_product(o.c, ptr(g.astype(np.uint64).reshape(-1).copy()), l, B, ptr(ct.copy()), ptr(exp))
    d = u64(2 * l * N)
    o.L.orc_decompose(o.c, ptr(ct.copy()), l, B, ptr(d))
    d = d.reshape(2 * l, N)
    acc = np.zeros((2, N), dtype=object)
    for poly in range(2):
        last = g[poly * l + l - 1]
        for j in range(l - 1):
            folded = (g[poly * l + j] - (2 ** ((l - 1 - j) * B)) * last) % q
            acc += o.ntt(d[poly * l + j]).astype(object) * folded
        acc += o.ntt(ct[poly * N:(poly + 1) * N]).astype(object) * last
    got = (acc % q).astype(np.uint64).reshape(-1)
    assert np.array_equal(got, exp)


def test_external_product_selects():
    o = Oracle()
    sk = o.keygen(41)
    msg = np.array([(i * 11) % o.t for i in range(o.N)], dtype=np.uint64)
    ct = o.encrypt(41, 9, sk, msg)
    for bit in (0, 1):
        g = o.rgsw(41, 3 + bit, sk, bit)
        out = u64(2 * o.N)
        o.L.orc_external_product(o.c, ptr(g), o.l, o.base_log2, ptr(ct), ptr(out))
        coeff = np.concatenate([o.intt(out[:o.N]), o.intt(out[o.N:])])
        assert np.array_equal(o.decrypt(sk, coeff), msg * bit)


# ---------------------------------------------------------------------------------
# server stages
# ---------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c1():
    o = Oracle(fst=256, other=32)
    sk = o.keygen(11)
    return dict(o=o, sk=sk, gk=o.galois_keys(11, sk), rk=o.rgsw_key(11, sk))


def test_query_indices_msb_first(c1):
    o = c1["o"]
    # plaintext index = row * fst + col; row bits MSB first (client.cpp:38-53)
    pt = 19 * 256 + 77
    assert list(map(int, o.query_indices(pt))) == [77, 1, 0, 0, 1, 1]


def test_expansion_decrypts_to_selection_vector(c1):
    o, sk = c1["o"], c1["sk"]
    pt = 19 * 256 + 77
    leaves = o.expand(o.query(11, 3, sk, pt), c1["gk"]).reshape(o.u, 2 * o.N)
    delta = o.q // o.t
    bits = o.query_indices(pt)[1:]
    gad = [pow(2, (o.l - 1 - k) * o.base_log2, o.q) for k in range(o.l)]
    for j in range(o.u):
        ph = o.phase(sk, leaves[j])
        if j < o.fst:
            want = delta if j == 77 else 0
        else:
            d, k = divmod(j - o.fst, o.l)
            want = int(bits[d]) * gad[k]
        err = (int(ph[0]) - want) % o.q
        err = min(err, o.q - err)
        assert err < 2**40, (j, err)
        rest = np.minimum(ph[1:], o.q - ph[1:])
        assert int(rest.max()) < 2**40


def test_first_dim_raw_brute_force():
    n, fst, other = 16, 4, 4
    o = Oracle(N=n, fst=fst, other=other, l=1, l_key=1)
    rng = random.Random(12)
    db = np.array([rng.randrange(o.q) for _ in range(n * fst * other)], dtype=np.uint64)
    sel = np.array([rng.randrange(o.q) for _ in range(fst * 2 * n)], dtype=np.uint64)
    inter = np.zeros(n * 2 * other * 2, dtype=np.uint64)  # uint128 as (lo, hi)
    o.L.orc_first_dim_raw(o.c, ptr(db), ptr(sel), ptr(inter))
    inter = inter.reshape(n, other, 2, 2)
    for s in range(n):
        for r in range(other):
            for c in range(2):
                want = sum(int(db[s * fst * other + r * fst + j]) * int(sel[j * 2 * n + c * n + s])
                           for j in range(fst)) % 2**128
                got = int(inter[s, r, c, 0]) | (int(inter[s, r, c, 1]) << 64)
                assert got == want


def test_first_dim_raw_wraps_mod_2_128():
    n, fst, other = 2048, 512, 2
    o = Oracle(N=n, fst=fst, other=other)
    db = np.full(n * fst * other, o.q - 1, dtype=np.uint64)
    sel = np.full(fst * 2 * n, o.q - 1, dtype=np.uint64)
    inter = np.zeros(n * 2 * other * 2, dtype=np.uint64)
    o.L.orc_first_dim_raw(o.c, ptr(db), ptr(sel), ptr(inter))
    full = fst * (o.q - 1) ** 2
    assert full >= 2**128
    got = int(inter[0]) | (int(inter[1]) << 64)
    assert got == full % 2**128


def test_mux_selects_half(c1):
    o, sk = c1["o"], c1["sk"]
    msgs = [np.full(o.N, 100 + i, dtype=np.uint64) for i in range(4)]
    cts = np.concatenate([o.encrypt(11, 50 + i, sk, m) for i, m in enumerate(msgs)])
    for bit in (0, 1):
        out = o.mux_level(cts, o.rgsw(11, 60 + bit, sk, bit)).reshape(2, 2 * o.N)
        for i in range(2):
            assert np.array_equal(o.decrypt(sk, out[i]), msgs[i + 2 * bit])


def x87_mod_switch(x, q_big, q_small):
    """x87 semantics in exact rationals: double scale, product rounded to a
    64-bit significand (ties to even), then round half away from zero."""
    s = Fraction(float(q_small) / float(q_big))
    v = x * s
    if v == 0:
        return 0
    e = v.numerator.bit_length() - v.denominator.bit_length()
    while Fraction(2) ** e > v:
        e -= 1
    while Fraction(2) ** (e + 1) <= v:
        e += 1
    ulp = Fraction(2) ** (e - 63)
    k, rem = divmod(v, ulp)
    if rem > ulp / 2 or (rem == ulp / 2 and k % 2 == 1):
        k += 1
    v = k * ulp
    fl = v.numerator // v.denominator
    return fl + 1 if v - fl >= Fraction(1, 2) else fl


def test_mod_switch_three_ways():
    o = Oracle()
    rng = random.Random(13)
    xs = [0, 1, 2, Q - 1, Q // 2, Q // 2 + 1, 2**33, 8589934591, 8589934593]
    # values whose exact product lands on or next to a rounding boundary
    ratio = Fraction(float(SMALL_Q) / float(Q))
    for k in (1, 2, 1000, 123456, SMALL_Q - 1):
        c = int((k + Fraction(1, 2)) / ratio)
        xs += [c - 1, c, c + 1]
    xs += [rng.randrange(Q) for _ in range(300)]
    for x in xs:
        a = o.L.orc_mod_switch_one(x, Q, SMALL_Q)
        b = o.L.orc_mod_switch_one_int(x, Q, SMALL_Q)
        assert a == b == x87_mod_switch(x, Q, SMALL_Q), x


def test_pack_layout_and_roundtrip():
    o = Oracle()
    rng = np.random.default_rng(14)
    ct = rng.integers(0, o.small_q + 1, size=2 * o.N, dtype=np.uint64)
    ct[0] = o.small_q
    data = o.pack(ct)
    assert len(data) == 2 * 2048 * 27 // 8 == 13824
    bits = int.from_bytes(data, "little")
    for i in (0, 1, 7, 8, 2047, 2048, 4095):
        assert (bits >> (27 * i)) & (2**27 - 1) == int(ct[i]) & (2**27 - 1)
    assert np.array_equal(o.unpack(data), ct & np.uint64(2**27 - 1))
    with pytest.raises(RuntimeError):
        o.unpack(data[:-1])


def test_db_entry_mode_layout():
    o = Oracle(fst=16, other=2)
    pt = o.db_plaintext(5, 1, 0x1234)
    entry = o.entry(pt)
    assert entry[:8] == (0x1234).to_bytes(8, "big")
    assert len(entry) == 4096
    assert int(pt.max()) < 2**16
    for i in (0, 3, 4, 1000):
        assert int(pt[i]) == entry[2 * i] | (entry[2 * i + 1] << 8)
    u = o.db_plaintext(5, 0, 7)
    assert int(u.max()) < o.t


def _balanced_lift(v, t, q):
    """The residue mod q of the balanced representative of v mod t, i.e. the
    integer in [-(t-1)/2, (t-1)/2] congruent to v mod t, read mod q."""
    half = (t - 1) // 2
    return ((v + half) % t - half) % q


def _check_plain_lift(o, lift_ntt, values):
    """lift_ntt(coeffs) -> NTT-form words.  The words must be the O(N^2)
    negacyclic evaluation (SEAL order, minimal root) of the balanced lift."""
    n, q, t = o.N, o.q, o.t
    logn = n.bit_length() - 1
    psi = min_root_python(q, n)
    w = [[pow(psi, (2 * bitrev(j, logn) + 1) * i, q) for i in range(n)] for j in range(n)]
    for k in range(0, len(values), n):
        chunk = [values[(k + i) % len(values)] for i in range(n)]
        lifted = [_balanced_lift(v, t, q) for v in chunk]
        for v, x in zip(chunk, lifted):
            signed = x if x < q // 2 else x - q
            if not (-(t - 1) // 2 <= signed <= (t - 1) // 2 and (signed - v) % t == 0):
                return False
        want = [sum(a * wj for a, wj in zip(lifted, w[j])) % q for j in range(n)]
        if list(map(int, lift_ntt(np.array(chunk, dtype=np.uint64)))) != want:
            return False
    return True


@pytest.mark.parametrize("plain_bits,n", [(16, 16), (17, 16), (16, 32), (17, 32)])
def test_plain_lift_is_balanced_representative(plain_bits, n):
    """SEAL transform_to_ntt_inplace(Plaintext) as PirServer::preprocess_ntt uses
    it (src/server.cpp:703-711): every v in [0, t) must land on the balanced
    representative of v mod t before the NTT.  Checked for all of [0, t) at
    t = 32771 and 65537, with the edges (t-1)/2 and (t+1)/2 in their own polynomial."""
    o = Oracle(N=n, plain_bits=plain_bits, fst=2, other=1, l=1, l_key=1)
    t, q = o.t, o.q
    assert t == {16: 32771, 17: 65537}[plain_bits]

    def oracle_lift(co):
        out = u64(n)
        o.L.orc_plain_to_ntt(o.c, ptr(np.ascontiguousarray(co)), ptr(out))
        return out

    half = (t - 1) // 2
    edges = [half, half + 1, 0, t - 1, half - 1, half + 2]
    assert _check_plain_lift(o, oracle_lift, edges)
    assert _check_plain_lift(o, oracle_lift, list(range(t)))

    # the check rejects the plausible slips: threshold t/2 (rounded down) and
    # increment q - t + 1, each restated around the oracle's own transform
    def slipped(thr, inc):
        def f(co):
            x = np.array([int(v) + inc if int(v) >= thr else int(v) for v in co], dtype=np.uint64)
            return o.ntt(x)
        return f

    assert _check_plain_lift(o, slipped((t + 1) // 2, q - t), edges)
    assert not _check_plain_lift(o, slipped(t // 2, q - t), edges)
    assert not _check_plain_lift(o, slipped((t + 1) // 2, q - t + 1), edges)


@pytest.mark.parametrize("mode", [0, 1])
def test_end_to_end_record_recovered(c1, mode):
    o, sk = c1["o"], c1["sk"]
    db = o.db_build(7, mode)
    rkn = o.rgsw_key_ntt(c1["rk"])
    for idx in (0, 3333, o.num_pt - 1):
        resp, _ = o.answer(db, c1["gk"], rkn, o.query(11, 200 + idx, sk, idx))
        assert int(resp.max()) <= o.small_q
        wire = o.unpack(o.pack(resp))
        pt = o.decrypt_small(sk, wire)
        assert np.array_equal(pt, o.db_plaintext(7, mode, idx))
        if mode == 1:
            assert o.entry(pt)[:8] == idx.to_bytes(8, "big")
