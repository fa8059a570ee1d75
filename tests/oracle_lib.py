"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# the portable build; ONION_ORACLE_LIB may name another build of the same
# sources (the benchmark's -march=native one)
LIB_PATH = os.environ.get("ONION_ORACLE_LIB") or os.path.join(ORACLE_DIR, "_build", "liboracle.so")

U64P = C.POINTER(C.c_uint64)


class OrcNtt(C.Structure):
    _fields_ = [("N", C.c_uint32), ("logN", C.c_uint32),
                ("mod_value", C.c_uint64), ("ratio_lo", C.c_uint64), ("ratio_hi", C.c_uint64),
                ("psi", C.c_uint64), ("rp", U64P), ("rp_shoup", U64P), ("irp", U64P),
                ("irp_shoup", U64P), ("n_inv", C.c_uint64), ("n_inv_shoup", C.c_uint64)]


class OrcMod(C.Structure):
    _fields_ = [("value", C.c_uint64), ("ratio_lo", C.c_uint64), ("ratio_hi", C.c_uint64)]


class OrcCtx(C.Structure):
    _fields_ = [("N", C.c_uint32), ("logN", C.c_uint32),
                ("q", C.c_uint64), ("p", C.c_uint64), ("t", C.c_uint64), ("small_q", C.c_uint64),
                ("plain_bits", C.c_uint32), ("l", C.c_uint32), ("base_log2", C.c_uint32),
                ("l_key", C.c_uint32), ("base_log2_key", C.c_uint32),
                ("fst", C.c_uint32), ("other", C.c_uint32), ("ndim", C.c_uint32),
                ("h", C.c_uint32), ("u", C.c_uint32), ("num_pt", C.c_uint64),
                ("mq", OrcMod), ("mp", OrcMod), ("msq", OrcMod),
                ("ntt_q", OrcNtt), ("ntt_p", OrcNtt), ("ntt_sq", OrcNtt)]


class OrcTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("expand", "convert", "first_dim", "first_dim_core", "other_dim", "mod_switch", "total")]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.environ.get("ONION_ORACLE_LIB") or LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER(OrcCtx)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_uint32] * 6
        L.orc_destroy.argtypes = [P]
        for name, res, args in [
            ("orc_is_prime", C.c_int, [C.c_uint64]),
            ("orc_get_primes", C.c_size_t, [C.c_uint64, C.c_uint32, C.c_size_t, U64P]),
            ("orc_generate_prime", C.c_uint64, [C.c_uint32]),
            ("orc_barrett128", C.c_uint64, [C.POINTER(OrcMod), C.c_uint64, C.c_uint64]),
            ("orc_bit_reverse", C.c_uint64, [C.c_uint64, C.c_uint32]),
            ("orc_rand64", C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            ("orc_ntt_fwd", None, [P, C.c_int, U64P]),
            ("orc_ntt_inv", None, [P, C.c_int, U64P]),
            ("orc_ntt_fwd_naive", None, [P, C.c_int, U64P, U64P]),
            ("orc_ntt_psi", C.c_uint64, [P, C.c_int]),
            ("orc_galois_coeff", None, [P, U64P, C.c_uint32, C.c_uint64, U64P]),
            ("orc_shift_neg", None, [P, U64P, C.c_uint64, U64P]),
            ("orc_keyswitch_add", None, [P, U64P, U64P, U64P]),
            ("orc_apply_galois", None, [P, U64P, C.c_uint32, U64P, U64P]),
            ("orc_decompose", None, [P, U64P, C.c_uint32, C.c_uint32, U64P]),
            ("orc_external_product", None, [P, U64P, C.c_uint32, C.c_uint32, U64P, U64P]),
            ("orc_expand", None, [P, U64P, U64P, U64P]),
            ("orc_convert", None, [P, U64P, U64P, U64P]),
            ("orc_rgsw_key_to_ntt", None, [P, U64P, U64P]),
            ("orc_first_dim_raw", None, [P, U64P, U64P, U64P]),
            ("orc_first_dim", None, [P, U64P, U64P, U64P, C.POINTER(OrcTimes)]),
            ("orc_mux_level", None, [P, U64P, C.c_uint64, U64P]),
            ("orc_mod_switch", None, [P, U64P]),
            ("orc_mod_switch_one", C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            ("orc_mod_switch_one_int", C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            ("orc_pack_response", C.c_size_t, [P, U64P, C.c_char_p]),
            ("orc_unpack_response", C.c_int, [P, C.c_char_p, C.c_size_t, U64P]),
            ("orc_answer", C.c_int, [P, U64P, U64P, U64P, U64P, U64P, C.POINTER(OrcTimes)]),
            ("orc_db_plaintext", None, [P, C.c_uint64, C.c_int, C.c_uint64, U64P]),
            ("orc_db_build", None, [P, C.c_uint64, C.c_int, U64P]),
            ("orc_db_build_range", None, [P, C.c_uint64, C.c_int, C.c_uint64, C.c_uint64, U64P]),
            ("orc_plain_to_ntt", None, [P, U64P, U64P]),
            ("orc_client_keygen", None, [P, C.c_uint64, U64P]),
            ("orc_client_encrypt_zero", None, [P, C.c_uint64, C.c_uint64, U64P, U64P]),
            ("orc_client_encrypt", None, [P, C.c_uint64, C.c_uint64, U64P, U64P, U64P]),
            ("orc_client_galois_keys", None, [P, C.c_uint64, U64P, U64P]),
            ("orc_client_rgsw_key", None, [P, C.c_uint64, U64P, U64P]),
            ("orc_client_rgsw", None, [P, C.c_uint64, C.c_uint64, U64P, C.c_uint64, U64P]),
            ("orc_client_query_indices", None, [P, C.c_uint64, U64P]),
            ("orc_client_query", None, [P, C.c_uint64, C.c_uint64, U64P, C.c_uint64, U64P]),
            ("orc_client_phase", None, [P, U64P, U64P, U64P]),
            ("orc_client_decrypt", None, [P, U64P, U64P, U64P]),
            ("orc_client_decrypt_small", None, [P, U64P, U64P, U64P]),
            ("orc_entry_from_plaintext", None, [P, U64P, C.c_char_p]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(U64P)


def u64(n) -> np.ndarray:
    return np.zeros(n, dtype=np.uint64)


class Oracle:
    """One oracle context (ring, moduli, NTT tables, database shape)."""

    def __init__(self, N=2048, plain_bits=17, l=5, l_key=10, fst=256, other=32):
        self.L = lib()
        self.c = self.L.orc_create(N, plain_bits, l, l_key, fst, other)
        if not self.c:
            raise ValueError("invalid oracle parameters")
        s = self.c.contents
        for f in ("N", "logN", "q", "p", "t", "small_q", "l", "base_log2", "l_key",
                  "base_log2_key", "fst", "other", "ndim", "h", "u", "num_pt", "plain_bits"):
            setattr(self, f, getattr(s, f))

    def __del__(self):
        try:
            self.L.orc_destroy(self.c)
        except Exception:
            pass

    # ---- arithmetic -------------------------------------------------------------
    def ntt(self, a, which=0):
        x = np.array(a, dtype=np.uint64)
        self.L.orc_ntt_fwd(self.c, which, ptr(x))
        return x

    def intt(self, a, which=0):
        x = np.array(a, dtype=np.uint64)
        self.L.orc_ntt_inv(self.c, which, ptr(x))
        return x

    def ntt_naive(self, a, which=0):
        x = np.array(a, dtype=np.uint64)
        out = u64(self.N)
        self.L.orc_ntt_fwd_naive(self.c, which, ptr(x), ptr(out))
        return out

    def barrett(self, which, x: int) -> int:
        m = [self.c.contents.mq, self.c.contents.mp, self.c.contents.msq][which]
        return self.L.orc_barrett128(C.byref(m), x & (2**64 - 1), x >> 64)

    # ---- client ------------------------------------------------------------------
    def keygen(self, seed):
        sk = u64(self.N)
        self.L.orc_client_keygen(self.c, seed, ptr(sk))
        return sk

    def galois_keys(self, seed, sk):
        gk = u64(self.h * 4 * self.N)
        self.L.orc_client_galois_keys(self.c, seed, ptr(sk), ptr(gk))
        return gk

    def rgsw_key(self, seed, sk):
        rk = u64(2 * self.l_key * 2 * self.N)
        self.L.orc_client_rgsw_key(self.c, seed, ptr(sk), ptr(rk))
        return rk

    def rgsw_key_ntt(self, rk):
        out = u64(rk.size)
        self.L.orc_rgsw_key_to_ntt(self.c, ptr(rk), ptr(out))
        return out

    def query(self, seed, nonce, sk, pt_index):
        ct = u64(2 * self.N)
        self.L.orc_client_query(self.c, seed, nonce, ptr(sk), pt_index, ptr(ct))
        return ct

    def encrypt(self, seed, nonce, sk, msg):
        ct = u64(2 * self.N)
        m = np.ascontiguousarray(msg, dtype=np.uint64)
        self.L.orc_client_encrypt(self.c, seed, nonce, ptr(sk), ptr(m), ptr(ct))
        return ct

    def decrypt(self, sk, ct):
        pt = u64(self.N)
        self.L.orc_client_decrypt(self.c, ptr(sk), ptr(np.ascontiguousarray(ct)), ptr(pt))
        return pt

    def phase(self, sk, ct):
        ph = u64(self.N)
        self.L.orc_client_phase(self.c, ptr(sk), ptr(np.ascontiguousarray(ct)), ptr(ph))
        return ph

    def decrypt_small(self, sk, resp):
        pt = u64(self.N)
        self.L.orc_client_decrypt_small(self.c, ptr(sk), ptr(np.ascontiguousarray(resp)), ptr(pt))
        return pt

    def rgsw(self, seed, nonce, sk, bit):
        g = u64(2 * self.l * 2 * self.N)
        self.L.orc_client_rgsw(self.c, seed, nonce, ptr(sk), bit, ptr(g))
        return g

    def query_indices(self, pt_index):
        idx = u64(self.ndim)
        self.L.orc_client_query_indices(self.c, pt_index, ptr(idx))
        return idx

    # ---- database ----------------------------------------------------------------
    def db_plaintext(self, seed, mode, pt):
        co = u64(self.N)
        self.L.orc_db_plaintext(self.c, seed, mode, pt, ptr(co))
        return co

    def plain_to_ntt(self, coeffs):
        out = u64(self.N)
        self.L.orc_plain_to_ntt(self.c, ptr(np.ascontiguousarray(coeffs, dtype=np.uint64)), ptr(out))
        return out

    def db_build(self, seed, mode=0):
        db = u64(self.num_pt * self.N)
        self.L.orc_db_build(self.c, seed, mode, ptr(db))
        return db

    # ---- server ------------------------------------------------------------------
    def expand(self, query, gk):
        leaves = u64(self.u * 2 * self.N)
        self.L.orc_expand(self.c, ptr(query), ptr(gk), ptr(leaves))
        return leaves

    def convert(self, leaves, rk_ntt):
        out = u64((self.ndim - 1) * 2 * self.l * 2 * self.N)
        self.L.orc_convert(self.c, ptr(leaves), ptr(rk_ntt), ptr(out))
        return out

    def first_dim(self, db, sel_coeff):
        out = u64(self.other * 2 * self.N)
        self.L.orc_first_dim(self.c, ptr(db), ptr(np.ascontiguousarray(sel_coeff)), ptr(out), None)
        return out

    def mux_level(self, cts, rgsw):
        x = np.array(cts, dtype=np.uint64).reshape(-1)
        n = x.size // (2 * self.N)
        self.L.orc_mux_level(self.c, ptr(x), n, ptr(np.ascontiguousarray(rgsw)))
        return x[: (n // 2) * 2 * self.N].copy()

    def answer(self, db, gk, rk_ntt, query):
        resp = u64(2 * self.N)
        tm = OrcTimes()
        self.L.orc_answer(self.c, ptr(db), ptr(gk), ptr(rk_ntt), ptr(query), ptr(resp), C.byref(tm))
        return resp, {f: getattr(tm, f) for f, _ in OrcTimes._fields_}

    def pack(self, ct):
        buf = C.create_string_buffer(2 * self.N * 8)
        n = self.L.orc_pack_response(self.c, ptr(np.ascontiguousarray(ct)), buf)
        return buf.raw[:n]

    def unpack(self, data: bytes):
        ct = u64(2 * self.N)
        rc = self.L.orc_unpack_response(self.c, data, len(data), ptr(ct))
        if rc != 0:
            raise RuntimeError("unexpected end of response stream")
        return ct

    def entry(self, pt):
        n = (self.plain_bits - 1) * self.N // 8
        buf = C.create_string_buffer(n)
        self.L.orc_entry_from_plaintext(self.c, ptr(np.ascontiguousarray(pt)), buf)
        return buf.raw
