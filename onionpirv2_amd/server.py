"""Host-side handle of the MI355X PIR server (thin layer over the C ABI).

``GpuServer`` owns one ``onion_server``: the NTT-form database resident in
HBM, the per-client key store and the per-query scratch.  Its methods map one
to one onto include/onionpir_gpu.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import OnionParamInfo, OnionParams, OnionStageTimes, check, u64ptr


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _derive_sizes(p: OnionParams) -> dict:
    info = OnionParamInfo()
    check(_capi.lib().onion_params_derive(C.byref(p), C.byref(info)))
    return {f: getattr(info, f) for f, _ in OnionParamInfo._fields_}


def derive_params(fst_dim: int, other_dim: int, poly_degree: int = 2048, plain_bits: int = 17,
                  gsw_l: int = 5, gsw_l_key: int = 10) -> dict:
    """Moduli, gadget bases and shape of a configuration (no device needed)."""
    return _derive_sizes(OnionParams(poly_degree, plain_bits, gsw_l, gsw_l_key, fst_dim, other_dim))


class GpuServer:
    def __init__(self, fst_dim: int, other_dim: int, poly_degree: int = 2048, plain_bits: int = 17,
                 gsw_l: int = 5, gsw_l_key: int = 10):
        L = _capi.lib()
        self._L = L
        p = OnionParams(poly_degree, plain_bits, gsw_l, gsw_l_key, fst_dim, other_dim)
        h = C.c_void_p()
        check(L.onion_gpu_create(C.byref(p), C.byref(h)))
        self._h = h
        info = OnionParamInfo()
        check(L.onion_gpu_param_info(self._h, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in OnionParamInfo._fields_}
        self.N = poly_degree
        self.fst_dim, self.other_dim = fst_dim, other_dim
        self.gsw_l, self.gsw_l_key = gsw_l, gsw_l_key

    # ---- lifetime -----------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.onion_gpu_destroy(self._h)
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

    # ---- database -------------------------------------------------------------------
    def db_generate(self, seed: int, mode: int = 0) -> None:
        check(self._L.onion_gpu_db_generate(self._h, seed, mode))

    def db_upload_plain(self, pt_lo: int, coeffs) -> None:
        a = _u64(coeffs).reshape(-1)
        n = a.size // self.N
        check(self._L.onion_gpu_db_upload_plain(self._h, pt_lo, n, u64ptr(a)))

    def db_read_ntt(self, pt: int) -> np.ndarray:
        out = np.zeros(self.N, np.uint64)
        check(self._L.onion_gpu_db_read_ntt(self._h, pt, u64ptr(out)))
        return out

    # ---- keys -------------------------------------------------------------------------
    def set_galois_key(self, client_id: int, keys) -> None:
        a = _u64(keys).reshape(-1)
        check(self._L.onion_gpu_set_galois_key(self._h, client_id, u64ptr(a), a.size))

    def set_rgsw_key(self, client_id: int, rows) -> None:
        a = _u64(rows).reshape(-1)
        check(self._L.onion_gpu_set_rgsw_key(self._h, client_id, u64ptr(a), a.size))

    def drop_client(self, client_id: int) -> None:
        check(self._L.onion_gpu_drop_client(self._h, client_id))

    def share_keys(self, other: "GpuServer") -> None:
        """Use other's device key store (this server must hold no keys yet)."""
        check(self._L.onion_gpu_share_keys(self._h, other._h))
        self._key_owner = other  # keep the store's other user alive

    # ---- answers ------------------------------------------------------------------------
    def answer(self, client_id: int, query) -> np.ndarray:
        q = _u64(query).reshape(-1)
        out = np.zeros(2 * self.N, np.uint64)
        check(self._L.onion_gpu_answer(self._h, client_id, u64ptr(q), u64ptr(out)))
        return out

    def answer_batch(self, client_ids, queries) -> np.ndarray:
        ids = _u64(client_ids).reshape(-1)
        q = _u64(queries).reshape(-1)
        out = np.zeros(ids.size * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_answer_batch(self._h, ids.size, u64ptr(ids), u64ptr(q), u64ptr(out)))
        return out.reshape(ids.size, 2 * self.N)

    def answer_batch_noalloc(self, client_ids, queries) -> np.ndarray:
        """answer_batch with its scratch in the library's static device array
        (no device allocation; the batch must fit the array)."""
        ids = _u64(client_ids).reshape(-1)
        q = _u64(queries).reshape(-1)
        out = np.zeros(ids.size * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_answer_batch_noalloc(self._h, ids.size, u64ptr(ids), u64ptr(q), u64ptr(out)))
        return out.reshape(ids.size, 2 * self.N)

    def set_lanes(self, n_lanes: int) -> None:
        check(self._L.onion_gpu_set_lanes(self._h, n_lanes))

    def answer_batch_lane(self, lane: int, client_ids, queries):
        """Batch on one lane; returns (responses, stage-time dict of this call)."""
        ids = _u64(client_ids).reshape(-1)
        q = _u64(queries).reshape(-1)
        out = np.zeros(ids.size * 2 * self.N, np.uint64)
        t = OnionStageTimes()
        check(self._L.onion_gpu_answer_batch_lane(self._h, lane, ids.size, u64ptr(ids), u64ptr(q),
                                                  u64ptr(out), C.byref(t)))
        return out.reshape(ids.size, 2 * self.N), {f: getattr(t, f) for f, _ in OnionStageTimes._fields_}

    def scratch_bytes_per_query(self) -> int:
        b = C.c_uint64()
        check(self._L.onion_gpu_scratch_bytes_per_query(self._h, C.byref(b)))
        return b.value

    def scratch_bytes(self, cap: int) -> int:
        """Device bytes of one lane's scratch for batches of up to cap queries."""
        b = C.c_uint64()
        check(self._L.onion_gpu_scratch_bytes(self._h, cap, C.byref(b)))
        return b.value

    def pack_response(self, resp) -> bytes:
        r = _u64(resp).reshape(-1)
        cap = self.info["response_bytes"]
        buf = C.create_string_buffer(cap)
        n = C.c_size_t()
        check(self._L.onion_gpu_pack_response(self._h, u64ptr(r), buf, cap, C.byref(n)))
        return buf.raw[: n.value]

    def get_tuning(self) -> dict:
        t = _capi.OnionTuning()
        check(self._L.onion_gpu_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.OnionTuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update(kw)
        t = _capi.OnionTuning(*[cur[f] for f, _ in _capi.OnionTuning._fields_])
        check(self._L.onion_gpu_set_tuning(self._h, C.byref(t)))

    def autotune(self, batch_sizes=(1, 64), reps: int = 3) -> dict:
        """Pick and apply the fastest launch shapes for these batch sizes."""
        sizes = (C.c_uint32 * len(batch_sizes))(*batch_sizes)
        t = _capi.OnionTuning()
        check(self._L.onion_gpu_autotune(self._h, sizes, len(batch_sizes), reps, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.OnionTuning._fields_}

    def stage_times(self) -> dict:
        t = OnionStageTimes()
        check(self._L.onion_gpu_last_stage_times(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in OnionStageTimes._fields_}

    # ---- taps ---------------------------------------------------------------------------
    def tap_expand(self, client_id: int, query) -> np.ndarray:
        q = _u64(query).reshape(-1)
        out = np.zeros(self.info["useful_cnt"] * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_expand(self._h, client_id, u64ptr(q), u64ptr(out)))
        return out

    def tap_convert(self, client_id: int, leaves) -> np.ndarray:
        lv = _u64(leaves).reshape(-1)
        nd = self.info["ndim"] - 1
        out = np.zeros(max(nd, 0) * 2 * self.gsw_l * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_convert(self._h, client_id, u64ptr(lv), u64ptr(out)))
        return out

    def tap_first_dim(self, sel) -> np.ndarray:
        s = _u64(sel).reshape(-1)
        out = np.zeros(self.other_dim * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_first_dim(self._h, u64ptr(s), u64ptr(out)))
        return out

    def tap_first_dim_batch(self, sel) -> np.ndarray:
        """sel: nq x fst x 2N selection cts -> nq x other x 2N."""
        s = _u64(sel).reshape(-1)
        nq = s.size // (self.fst_dim * 2 * self.N)
        out = np.zeros(nq * self.other_dim * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_first_dim_batch(self._h, nq, u64ptr(s), u64ptr(out)))
        return out.reshape(nq, -1)

    def tap_mux(self, cts, rgsw) -> np.ndarray:
        c = _u64(cts).reshape(-1)
        g = _u64(rgsw).reshape(-1)
        n = c.size // (2 * self.N)
        out = np.zeros((n // 2) * 2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_mux(self._h, u64ptr(c), n, u64ptr(g), u64ptr(out)))
        return out

    def tap_ntt(self, polys, which: int = 0, inverse: bool = False) -> np.ndarray:
        a = np.array(polys, dtype=np.uint64).reshape(-1)
        check(self._L.onion_gpu_tap_ntt(self._h, which, int(inverse), u64ptr(a), a.size // self.N))
        return a

    def tap_galois(self, client_id: int, level: int, ct) -> np.ndarray:
        c = _u64(ct).reshape(-1)
        out = np.zeros(2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_galois(self._h, client_id, level, u64ptr(c), u64ptr(out)))
        return out

    def tap_copy(self, n_bytes: int) -> int:
        out = C.c_uint64()
        check(self._L.onion_gpu_tap_copy(self._h, n_bytes, C.byref(out)))
        return out.value

    def tap_valu_probe(self, blocks: int, rounds: int) -> float:
        """Kernel time (ms) of the VALU issue-rate probe (blocks x 256 threads)."""
        ms = C.c_double()
        check(self._L.onion_gpu_tap_valu_probe(self._h, blocks, rounds, C.byref(ms)))
        return ms.value

    def tap_mod_switch(self, ct) -> np.ndarray:
        c = _u64(ct).reshape(-1)
        out = np.zeros(2 * self.N, np.uint64)
        check(self._L.onion_gpu_tap_mod_switch(self._h, u64ptr(c), u64ptr(out)))
        return out
