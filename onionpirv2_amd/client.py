"""Seeded PIR client (CPU) -- ctypes binding of include/onionpir_client.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import OnionError, OnionParams, OnionValueError, U64P, u64ptr

_SIGS = {
    "onion_client_last_error": (C.c_char_p, []),
    "onion_client_create": (C.c_int, [C.POINTER(OnionParams), C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_void_p)]),
    "onion_client_create_secure": (C.c_int, [C.POINTER(OnionParams), C.c_uint64, C.POINTER(C.c_void_p)]),
    "onion_client_destroy": (C.c_int, [C.c_void_p]),
    "onion_client_id": (C.c_int, [C.c_void_p, U64P]),
    "onion_client_galois_keys": (C.c_int, [C.c_void_p, U64P, C.c_size_t]),
    "onion_client_rgsw_key": (C.c_int, [C.c_void_p, U64P, C.c_size_t]),
    "onion_client_query": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, U64P]),
    "onion_client_unpack_response": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, U64P]),
    "onion_client_decrypt": (C.c_int, [C.c_void_p, U64P, U64P]),
    "onion_client_entry": (C.c_int, [C.c_void_p, C.c_uint64, U64P, C.c_char_p, C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "onion_chacha20_block": (None, [C.POINTER(C.c_uint32)] * 3),
}

_bound = False


def _lib():
    global _bound
    L = _capi.lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib().onion_client_last_error().decode(errors="replace")
        if rc == -1:
            raise OnionValueError(rc, msg)
        raise OnionError(rc, msg)


class GpuClient:
    """One client: secret key, Galois keys, RGSW(s) key, queries, decryption.

    seed=None (the default) draws every secret from ChaCha20 keyed by the
    operating system's entropy.  An integer seed gives a reproducible client
    for tests and benchmarks; it is NOT SECURE (64 bits of key, no
    cryptographic generator).
    """

    def __init__(self, fst_dim: int, other_dim: int, client_id: int, seed: int | None = None,
                 poly_degree: int = 2048, plain_bits: int = 17, gsw_l: int = 5, gsw_l_key: int = 10):
        L = _lib()
        self._L = L
        p = OnionParams(poly_degree, plain_bits, gsw_l, gsw_l_key, fst_dim, other_dim)
        h = C.c_void_p()
        if seed is None:
            _check(L.onion_client_create_secure(C.byref(p), client_id, C.byref(h)))
        else:
            _check(L.onion_client_create(C.byref(p), client_id, seed, C.byref(h)))
        self._h = h
        self._next_nonce = 0
        self.client_id = client_id
        self.N = poly_degree
        self.plain_bits = plain_bits
        # sizes follow the server's derivation
        from .server import _derive_sizes
        self.sizes = _derive_sizes(p)

    def close(self):
        if getattr(self, "_h", None):
            self._L.onion_client_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def galois_keys(self) -> np.ndarray:
        out = np.zeros(self.sizes["galois_key_words"], np.uint64)
        _check(self._L.onion_client_galois_keys(self._h, u64ptr(out), out.size))
        return out

    def rgsw_key(self) -> np.ndarray:
        out = np.zeros(self.sizes["rgsw_key_words"], np.uint64)
        _check(self._L.onion_client_rgsw_key(self._h, u64ptr(out), out.size))
        return out

    def query(self, pt_index: int, nonce: int | None = None) -> np.ndarray:
        """Encrypted query for a plaintext index.  Each query needs a nonce this
        client never used (< 2^48): by default the next value of a per-client
        counter; an explicit nonce that was used before is refused."""
        out = np.zeros(2 * self.N, np.uint64)
        if nonce is None:
            while True:
                nonce, self._next_nonce = self._next_nonce, self._next_nonce + 1
                rc = self._L.onion_client_query(self._h, pt_index, nonce, u64ptr(out))
                if rc != -1 or b"already used" not in self._L.onion_client_last_error():
                    break
            _check(rc)
            return out
        _check(self._L.onion_client_query(self._h, pt_index, nonce, u64ptr(out)))
        return out

    def unpack_response(self, data: bytes) -> np.ndarray:
        out = np.zeros(2 * self.N, np.uint64)
        _check(self._L.onion_client_unpack_response(self._h, data, len(data), u64ptr(out)))
        return out

    def decrypt(self, resp) -> np.ndarray:
        r = np.ascontiguousarray(resp, dtype=np.uint64)
        out = np.zeros(self.N, np.uint64)
        _check(self._L.onion_client_decrypt(self._h, u64ptr(r), u64ptr(out)))
        return out

    def entry(self, entry_index: int, pt) -> bytes:
        p = np.ascontiguousarray(pt, dtype=np.uint64)
        cap = (self.plain_bits - 1) * self.N // 8
        buf = C.create_string_buffer(cap)
        n = C.c_size_t()
        _check(self._L.onion_client_entry(self._h, entry_index, u64ptr(p), buf, cap, C.byref(n)))
        return buf.raw[: n.value]


def chacha20_block(key_words, ctr_nonce_words) -> list[int]:
    """The secure generator's ChaCha20 block (16 output words)."""
    k = (C.c_uint32 * 8)(*key_words)
    cn = (C.c_uint32 * 4)(*ctr_nonce_words)
    out = (C.c_uint32 * 16)()
    _lib().onion_chacha20_block(k, cn, out)
    return list(out)
