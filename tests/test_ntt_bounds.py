"""Bound argument of the device transforms' lazy arithmetic (CPU).

onionpirv2_amd/csrc/ntt.hpp keeps butterfly values unreduced: the twiddle
product mul_shoup_4q (arith.hpp) truncates the Shoup quotient to
x1 s1 + hi(x1 s0) + hi(x0 s1), so it lands in [0, 4q); forward stages add up
to 4q, and before stages 3, 5, 7 and 9 the input half loses 8q when its bit 63
is set (reduce_2p63: q = 2^60 - c, so 8q = 2^63 - 8c); the inverse's sum half
is (s mod 2^62) + (s >> 62) 4c (reduce_2p62: 4q = 2^62 - 4c), below
2^62 + 12c, and its difference half a twiddle product of u + 8q - v.  The
forward output is brought to [0, q) by ntt_fwd_final (2^60 = c
mod q, then one conditional subtraction) or below 2^63 by ntt_fwd_below_2p63.
This models those steps with Python integers and checks, for both 60-bit
primes on worst-case and random inputs, that the quotient is at most 2 below
the exact one, that every value stays below 2^64 and within the stated bounds
(2^63 + 8q forward, 2^62 + 12c inverse), that the final reductions are canonical or
below 2^63, and that the results agree mod q with exact arithmetic.  A
reduction step dropped from the schedule or a wrong bound in the butterflies
fails it.  The same holds for reduce128 (arith.hpp), which takes the 128-bit
multiply-accumulate sums to their canonical residue through 2^60 = c mod q.
"""
import random

import pytest

M64 = (1 << 64) - 1
N = 2048
PRIMES = (1152921504606748673, 1152921504606830593)  # q, p


def shoup4(x, w, ws, q):
    x0, x1 = x & 0xFFFFFFFF, x >> 32
    s0, s1 = ws & 0xFFFFFFFF, ws >> 32
    qh = x1 * s1 + ((x1 * s0) >> 32) + ((x0 * s1) >> 32)
    exact = (x * ws) >> 64
    assert exact - 2 <= qh <= exact
    r = (x * w - qh * q) & M64
    assert r == x * w - qh * q and r < 4 * q
    return r


def reduce_2p63(v, q):
    c8 = (1 << 63) - 8 * q
    assert 0 < c8 < 1 << 32
    r = (v >> 63) * c8 + (v & ((1 << 63) - 1))
    assert r == (v - 8 * q if v >= 1 << 63 else v) and r < 1 << 63
    return r


def reduce_2p62(v, q):
    c4 = (1 << 62) - 4 * q
    assert 0 < c4 < 1 << 32 and v <= M64
    r = (v >> 62) * c4 + (v & ((1 << 62) - 1))
    assert r % q == v % q and r < (1 << 62) + 3 * c4
    return r


def fwd_final(v, q):
    c = (1 << 60) - q
    assert 0 < c < 1 << 28
    r = (v >> 60) * c + (v & ((1 << 60) - 1))
    assert r < 2 * q
    return r - q if r >= q else r


def shoup_exact(x, w, ws, q):
    r = (x * w - ((x * ws) >> 64) * q) & M64
    assert r < 2 * q
    return r


def forward(a, q, tab, tabs, small, corrections=(3, 5, 7, 9)):
    x = list(a)
    for s in range(11):
        H = 1024 >> s
        for i in range(N):
            if i & H:
                continue
            m = N // (2 * H) + (i >> (11 - s))
            w, ws = tab[m], tabs[m]
            u = x[i]
            if s in corrections:
                u = reduce_2p63(u, q)
            if small and s == 0:  # ct_bfly_small: exact quotient for inputs below 2^32
                v = shoup_exact(x[i + H], w, ws, q)
                x[i], x[i + H] = u + v, u + 2 * q - v
            else:
                v = shoup4(x[i + H], w, ws, q)
                x[i], x[i + H] = u + v, u + 4 * q - v
            assert x[i] <= M64 and x[i + H] <= M64
        assert max(x) < (1 << 63) + 8 * q, f"stage {s}: {max(x) / q:.2f} q"
    return x


def inverse(a, q, tab, tabs, reduce=True):
    x = list(a)
    for s in range(11):
        H = 1 << s
        for i in range(N):
            if i & H:
                continue
            m = N // (2 * H) + (i >> (s + 1))
            u, v = x[i], x[i + H]
            assert u + v <= M64 and v < 8 * q and u + 8 * q <= M64
            x[i], x[i + H] = reduce_2p62(u + v, q) if reduce else u + v, shoup4(u + 8 * q - v, tab[m], tabs[m], q)
        assert max(x) < (1 << 62) + 12 * ((1 << 60) - q)
    return x


def exact_forward(a, q, tab):
    x = [v % q for v in a]
    for s in range(11):
        H = 1024 >> s
        for i in range(N):
            if not i & H:
                m = N // (2 * H) + (i >> (11 - s))
                v = x[i + H] * tab[m] % q
                x[i], x[i + H] = (x[i] + v) % q, (x[i] - v) % q
    return x


def exact_inverse(a, q, tab):
    x = [v % q for v in a]
    for s in range(11):
        H = 1 << s
        for i in range(N):
            if not i & H:
                m = N // (2 * H) + (i >> (s + 1))
                x[i], x[i + H] = (x[i] + x[i + H]) % q, (x[i] - x[i + H]) * tab[m] % q
    return x


@pytest.mark.parametrize("q", PRIMES)
def test_lazy_forward_and_inverse_bounds(q):
    rng = random.Random(q & 0xFFFF)
    tab = [rng.randrange(q) for _ in range(N)]  # any twiddles: the bounds hold for every w < q
    tabs = [(w << 64) // q for w in tab]
    cases = [([4 * q - 1] * N, False), ([(1 << 32) - 1] * N, True),
             ([rng.randrange(4 * q) for _ in range(N)], False), ([rng.randrange(1 << 32) for _ in range(N)], True)]
    for a, small in cases:
        out = forward(a, q, tab, tabs, small)
        ref = exact_forward(a, q, tab)
        assert [v % q for v in out] == ref
        assert [fwd_final(v, q) for v in out] == ref
        assert all(reduce_2p63(v, q) % q == r for v, r in zip(out, ref))
    for a in ([8 * q - 1] * N, [4 * q - 1] * N, [rng.randrange(8 * q) for _ in range(N)]):
        out = inverse(a, q, tab, tabs)
        ref = exact_inverse(a, q, tab)
        assert [v % q for v in out] == ref
        assert [fwd_final(v, q) for v in out] == ref


def test_final_reductions_on_extreme_values():
    for q in PRIMES:
        for v in (0, q - 1, q, 2 * q - 1, 8 * q - 1, 8 * q, (1 << 63) - 1, 1 << 63, (1 << 63) + 8 * q - 1,
                  (1 << 64) - 1):
            assert fwd_final(v, q) == v % q
            if v < (1 << 63) + 8 * q:
                assert reduce_2p63(v, q) % q == v % q


def test_a_missing_reduction_overflows_the_bound():
    """Without the reduction before stage 9 the worst case passes 2^63 + 8q."""
    q = PRIMES[0]
    rng = random.Random(5)
    tab = [rng.randrange(q) for _ in range(N)]
    tabs = [(w << 64) // q for w in tab]
    with pytest.raises(AssertionError):
        forward([4 * q - 1] * N, q, tab, tabs, False, corrections=(3, 5, 7))


def test_inverse_without_the_sum_reduction_overflows():
    """Without reduce_2p62 the sum half doubles each stage and passes 8q."""
    q = PRIMES[1]
    rng = random.Random(6)
    tab = [rng.randrange(q) for _ in range(N)]
    tabs = [(w << 64) // q for w in tab]
    with pytest.raises(AssertionError):
        inverse([4 * q - 1] * N, q, tab, tabs, reduce=False)


def reduce128(v, q, drop_h2=False):
    """Word-level model of reduce128 (arith.hpp): the canonical residue of a
    128-bit sum mod q = 2^60 - c, c < 2^20, through 2^60 = c mod q."""
    assert 0 <= v < 1 << 128
    c = (1 << 60) - q
    assert 0 < c < 1 << 20
    w0, w1, w2, w3 = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(4)]
    h0 = ((w2 << 32 | w1) >> 28) & 0xFFFFFFFF  # v_alignbit_b32
    h1 = ((w3 << 32 | w2) >> 28) & 0xFFFFFFFF
    h2 = 0 if drop_h2 else w3 >> 28
    a = h0 * c + ((w1 & 0x0FFFFFFF) << 32 | w0)
    assert a < 1 << 61
    b = h1 * c + (a >> 32) + ((h2 * c) << 32)
    assert h2 * c < 1 << 32 and b < 1 << 57
    t0 = (b & 0x0FFFFFFF) << 32 | (a & 0xFFFFFFFF)
    t1 = (b >> 28) & 0xFFFFFFFF
    assert t1 == b >> 28
    r = t1 * c + t0
    assert r < 2 * q and r <= M64
    return r - q if r >= q else r


def test_reduce128_is_the_canonical_residue():
    """The 128-bit multiply-accumulate sums (external products, the first
    dimension's C mod 2^128) are reduced through 2^60 = c mod q: every value
    below 2^128 gives v mod q, for both primes and for the largest c allowed."""
    rng = random.Random(11)
    moduli = list(PRIMES) + [(1 << 60) - ((1 << 20) - 1), (1 << 60) - 1]
    for q in moduli:
        edges = [0, 1, q - 1, q, 2 * q - 1, 2 * q, (1 << 60) - 1, 1 << 60, (1 << 64) - 1, 1 << 64,
                 (1 << 124) - 1, 1 << 124, (1 << 127) - 1, 1 << 127, (1 << 128) - 1,
                 ((1 << 128) - 1) // q * q, ((1 << 128) - 1) // q * q - 1, 20 * (8 * q - 1) * (q - 1)]
        vals = edges + [rng.getrandbits(128) for _ in range(4000)] + \
            [rng.getrandbits(rng.randrange(1, 129)) for _ in range(4000)]
        for v in vals:
            assert reduce128(v, q) == v % q, (q, v)


def test_reduce128_without_the_top_bits_fails():
    """Dropping h2 (bits 124..127) gives wrong residues for large sums."""
    q = PRIMES[0]
    bad = [v for v in (1 << 127, (1 << 128) - 1, 3 << 125) if reduce128(v, q, drop_h2=True) != v % q]
    assert bad
