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
fails it.
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
