"""The C-ABI library loads (no GPU needed) and exports every function that
include/*.h declares; host-side parameter derivation matches the oracle."""
import ctypes as C
import os
import re

import pytest

from oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(onion_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    from onionpirv2_amd import _capi
    lib = C.CDLL(_capi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from onionpirv2_amd import _capi, client, serve
    bound = set(_capi.exported_symbols()) | set(client._SIGS) | set(serve._SIGS)
    assert set(declared_functions()) <= bound


@pytest.mark.parametrize("fst,other", [(256, 32), (256, 512), (512, 1024), (128, 256), (512, 8192), (12, 4)])
def test_param_derivation_matches_oracle(fst, other):
    from onionpirv2_amd.server import derive_params
    d = derive_params(fst, other)
    o = Oracle(fst=fst, other=other)
    assert (d["q"], d["p"], d["t"], d["small_q"]) == (o.q, o.p, o.t, o.small_q)
    assert (d["base_log2"], d["base_log2_key"], d["ndim"], d["expan_height"], d["useful_cnt"],
            d["num_pt"]) == (o.base_log2, o.base_log2_key, o.ndim, o.h, o.u, o.num_pt)
    assert d["response_bytes"] == 13824


def test_invalid_shapes_are_rejected():
    from onionpirv2_amd import OnionValueError
    from onionpirv2_amd.server import derive_params
    with pytest.raises(OnionValueError):
        derive_params(256, 48)  # other_dim not a power of two
    with pytest.raises(OnionValueError):
        derive_params(2100, 4)  # 2^ceil(log2(2110)) > N: expansion height above log2 N
    # 2 l products of a transformed digit (< 8q) and a gadget value (< q) sum in
    # 128 bits: 32 stay below 256 q^2 < 2^128, 34 may not
    assert derive_params(64, 4, gsw_l=16, gsw_l_key=16)["base_log2_key"] == 4
    with pytest.raises(OnionValueError):
        derive_params(64, 4, gsw_l_key=17)
    with pytest.raises(OnionValueError):
        derive_params(64, 4, gsw_l=17)
    o = Oracle(fst=64, other=4)
    assert 256 * o.q * o.q < 1 << 128 < 288 * o.q * o.q
