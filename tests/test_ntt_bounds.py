"""Bound argument of the device transforms' lazy arithmetic (CPU).

onionpirv2_amd/csrc/ntt.hpp keeps butterfly values unreduced: the twiddle
product mul_shoup_4q (arith.hpp) truncates the Shoup quotient to
x1 s1 + hi(x1 s0) + hi(x0 s1), so it lands in [0, 4q); forward stages add up
to 4q and reduce the input half by 8q before stages 3, 5, 7 and 9; inverse
values stay below 4q.  This models those butterflies with Python integers and
checks, for both 60-bit primes on worst-case and random inputs, that the
quotient is at most 2 below the exact one, that every value stays below 2^64
and within the stated bounds (16q forward, 4q inverse), and that the results
agree mod q with exact arithmetic.  A reduction step dropped from the schedule
or a wrong bound in the butterflies fails it.
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
            if s in corrections and u >= 8 * q:
                u -= 8 * q
            if small and s == 0:  # ct_bfly_small: exact quotient for inputs below 2^32
                v = shoup_exact(x[i + H], w, ws, q)
                x[i], x[i + H] = u + v, u + 2 * q - v
            else:
                v = shoup4(x[i + H], w, ws, q)
                x[i], x[i + H] = u + v, u + 4 * q - v
            assert x[i] <= M64 and x[i + H] <= M64
        assert max(x) < 16 * q, f"stage {s}: {max(x) / q:.2f} q"
    return x


def inverse(a, q, tab, tabs):
    x = list(a)
    for s in range(11):
        H = 1 << s
        for i in range(N):
            if i & H:
                continue
            m = N // (2 * H) + (i >> (s + 1))
            u, v = x[i], x[i + H]
            sm = u + v
            if sm >= 4 * q:
                sm -= 4 * q
            x[i], x[i + H] = sm, shoup4(u + 4 * q - v, tab[m], tabs[m], q)
        assert max(x) < 4 * q
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
        assert [v % q for v in out] == exact_forward(a, q, tab)
    for a in ([4 * q - 1] * N, [rng.randrange(4 * q) for _ in range(N)]):
        out = inverse(a, q, tab, tabs)
        assert [v % q for v in out] == exact_inverse(a, q, tab)


def test_a_missing_reduction_overflows_the_bound():
    """Without the reduction before stage 9 the worst case passes 16q."""
    q = PRIMES[0]
    rng = random.Random(5)
    tab = [rng.randrange(q) for _ in range(N)]
    tabs = [(w << 64) // q for w in tab]
    with pytest.raises(AssertionError):
        forward([4 * q - 1] * N, q, tab, tabs, False, corrections=(3, 5, 7))
