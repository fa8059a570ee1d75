"""Host restatement of the device's seeded database generator (k_db_generate),
used to check retrieved records without reading the database back.

mode 0: coefficient e of plaintext pt = floor(r * t / 2^64),
        r = rand64(seed, stream 0x4442, pt*N + e)
mode 1: entry bytes -- 8-byte big-endian index, then bytes of
        rand64(seed, 0x4442, pt*N + b/8); bit_count(t)-1 bits per coefficient.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = 0x9E3779B97F4A7C15
_C = 0x632BE59BD9B4E019
STREAM_DB = 0x4442
MASK = (1 << 64) - 1


def _mix_int(z: int) -> int:
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & MASK
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & MASK
    z ^= z >> 31
    return z


def _mix_np(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = z ^ (z >> np.uint64(30))
        z = z * _M1
        z = z ^ (z >> np.uint64(27))
        z = z * _M2
        return z ^ (z >> np.uint64(31))


def stream_key(seed: int, stream: int) -> int:
    return _mix_int((seed * _G + _mix_int(stream ^ _C)) & MASK)


def rand64(key: int, ctr: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = np.uint64(key) + (ctr.astype(np.uint64) + np.uint64(1)) * np.uint64(_G)
    return _mix_np(x)


def plaintext(seed: int, mode: int, pt: int, t: int = 65537, N: int = 2048) -> np.ndarray:
    key = stream_key(seed, STREAM_DB)
    if mode == 0:
        r = rand64(key, np.arange(N, dtype=np.uint64) + np.uint64(pt * N))
        return np.array([(int(v) * t) >> 64 for v in r], dtype=np.uint64)
    bpc = t.bit_length() - 1
    nbytes = bpc * N // 8
    r = rand64(key, np.arange((nbytes + 7) // 8, dtype=np.uint64) + np.uint64(pt * N))
    b = r.view(np.uint8)[:nbytes].copy()
    b[:8] = np.frombuffer(int(pt).to_bytes(8, "big"), dtype=np.uint8)
    bits = int.from_bytes(b.tobytes(), "little")
    mask = (1 << bpc) - 1
    return np.array([(bits >> (bpc * i)) & mask for i in range(N)], dtype=np.uint64)
