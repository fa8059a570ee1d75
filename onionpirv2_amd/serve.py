"""Serving loop over device-resident databases (binding of include/onionpir_serve.h).

``GpuServe`` queues queries per database; a pool of worker threads in
libonionpir.so serves all databases: a free worker picks the database whose
oldest waiting query is the oldest of all and takes its waiting queries (up to
the database's admitted batch cap) as one batch on a free lane of that server.
Responses are bit-identical to ``GpuServer.answer`` on the same inputs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import ERRORS, U64P, OnionError, OnionKeyError, OnionValueError, u64ptr

ETIMEDOUT = -6
FOREVER = 0xFFFFFFFF


class ServeConfig(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("lanes_per_db", C.c_uint32), ("scratch_budget", C.c_uint64),
                ("linger_us", C.c_uint32), ("workers", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("mean_batch", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_ms", C.c_double),
                ("span_s", C.c_double), ("queries_per_s", C.c_double), ("batch_cap", C.c_uint32)]


_SIGS = {
    "onion_serve_last_error": (C.c_char_p, []),
    "onion_serve_create": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]),
    "onion_serve_destroy": (C.c_int, [C.c_void_p]),
    "onion_serve_add_db": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "onion_serve_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, U64P, C.POINTER(C.c_uint64)]),
    "onion_serve_wait": (C.c_int, [C.c_void_p, C.c_uint64, U64P, C.c_uint32]),
    "onion_serve_discard": (C.c_int, [C.c_void_p, C.c_uint64]),
    "onion_serve_stats_get": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(ServeStats)]),
    "onion_serve_reset_stats": (C.c_int, [C.c_void_p]),
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


class ServeTimeout(OnionError, TimeoutError):
    """The response is not ready yet."""


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = _lib().onion_serve_last_error().decode(errors="replace")
    if rc == ETIMEDOUT:
        raise ServeTimeout(rc, msg)
    if rc == -2:
        raise OnionKeyError(rc, msg)
    if rc == -1:
        raise OnionValueError(rc, msg)
    raise OnionError(rc, msg if rc in ERRORS else f"code {rc}: {msg}")


class GpuServe:
    def __init__(self, max_batch: int = 64, lanes_per_db: int = 2, scratch_budget: int = 0, linger_us: int = 0,
                 workers: int = 0):
        """workers: batches in flight over all databases (0 = one per lane of
        every database added)."""
        L = _lib()
        self._L = L
        cfg = ServeConfig(max_batch, lanes_per_db, scratch_budget, linger_us, workers)
        h = C.c_void_p()
        check(L.onion_serve_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._servers = []  # keep the servers alive while served
        self._ct_words = []

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.onion_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_db(self, server) -> tuple[int, int]:
        """Serve a GpuServer; returns (db_id, admitted batch cap)."""
        db, cap = C.c_uint32(), C.c_uint32()
        check(self._L.onion_serve_add_db(self._h, server._h, C.byref(db), C.byref(cap)))
        self._servers.append(server)
        self._ct_words.append(2 * server.N)
        return db.value, cap.value

    def submit(self, db_id: int, client_id: int, query) -> int:
        q = np.ascontiguousarray(query, dtype=np.uint64).reshape(-1)
        if q.size != self._ct_words[db_id]:
            raise OnionValueError(-1, f"query has {q.size} words, expected {self._ct_words[db_id]}")
        t = C.c_uint64()
        check(self._L.onion_serve_submit(self._h, db_id, client_id, u64ptr(q), C.byref(t)))
        return t.value

    def wait(self, ticket: int, db_id: int = 0, timeout_ms: int = FOREVER) -> np.ndarray:
        out = np.zeros(self._ct_words[db_id], np.uint64)
        check(self._L.onion_serve_wait(self._h, ticket, u64ptr(out), timeout_ms))
        return out

    def discard(self, ticket: int) -> None:
        """Retire a ticket without its response."""
        check(self._L.onion_serve_discard(self._h, ticket))

    def stats(self, db_id: int | None = None) -> dict:
        s = ServeStats()
        check(self._L.onion_serve_stats_get(self._h, 0xFFFFFFFF if db_id is None else db_id, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServeStats._fields_}

    def reset_stats(self) -> None:
        check(self._L.onion_serve_reset_stats(self._h))
