"""Which LDS exchanges of the device transforms stay inside one wave (CPU).

onionpirv2_amd/csrc/ntt.hpp runs a 2048-point transform on 256 threads (four
waves of 64) in four register passes.  Between passes thread t writes its 8
coefficients to LDS and reads the 8 of the next pass.  An exchange in which
every coefficient a wave reads was written by the same wave needs no
workgroup barrier: a wave's LDS operations complete in issue order, so the
transforms synchronise those exchanges per wave and keep workgroup barriers
for the one that crosses waves (pass A, stride 256).  This checks the
thread-to-coefficient maps the kernels use (pass_base) for both directions,
and that the stride-256 exchange does cross waves (so the check can fail).
"""
import pytest

N, T, WAVE = 2048, 256, 64


def pass_base(t, s0):
    return (t & ((1 << s0) - 1)) | ((t >> s0) << (s0 + 3))


# the passes as (base(t), stride): natural/A, B, C, D (blocked)
PASSES = {
    "A": (lambda t: t, 256),
    "B": (lambda t: pass_base(t, 5), 32),
    "C": (lambda t: pass_base(t, 2), 4),
    "D": (lambda t: 8 * t, 1),
}


def owned(p, t):
    base, stride = PASSES[p]
    return {base(t) + j * stride for j in range(8)}


def test_every_pass_is_a_partition():
    for p in PASSES:
        seen = []
        for t in range(T):
            seen += sorted(owned(p, t))
        assert sorted(seen) == list(range(N)), p


def wave_local(src, dst):
    for w in range(T // WAVE):
        wrote = set().union(*(owned(src, t) for t in range(w * WAVE, (w + 1) * WAVE)))
        read = set().union(*(owned(dst, t) for t in range(w * WAVE, (w + 1) * WAVE)))
        if read != wrote:
            return False
    return True


@pytest.mark.parametrize("src,dst", [("B", "C"), ("C", "D"), ("D", "C"), ("C", "B")])
def test_inner_exchanges_stay_in_one_wave(src, dst):
    assert wave_local(src, dst)


@pytest.mark.parametrize("src,dst", [("A", "B"), ("B", "A")])
def test_the_stride_256_exchange_crosses_waves(src, dst):
    assert not wave_local(src, dst)


def test_waves_use_disjoint_padded_lds_ranges():
    """The padded LDS index i + i/16 keeps each wave's coefficients
    [512 w, 512 w + 512) in a range no other wave touches."""
    pad = lambda i: i + (i >> 4)
    spans = [(pad(512 * w), pad(512 * w + 511)) for w in range(4)]
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert a1 < b0
