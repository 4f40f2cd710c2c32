"""The C-ABI library loads and exports every entry point include/cucm.h
declares; header-only helpers work without a GPU."""
import ctypes
import os
import re

import pytest

import cucm_amd
from cucm_amd import _native

HDR = os.path.join(os.path.dirname(__file__), "..", "include", "cucm.h")


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cucm_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_all_exported():
    L = cucm_amd.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_native.EXPORTS) | set(names)


def test_python_binding_lists_every_declared_symbol():
    assert set(declared_functions()) == set(_native.EXPORTS)


def test_params_struct_layout():
    assert ctypes.sizeof(_native.Params) == 20
    p = cucm_amd.make_params()
    assert (p.chunk_size, p.unit_bits, p.batch_chunks, p.table_bits) == (1 << 20, 128, 100, 25)


def test_compress_bound_is_strict():
    p = cucm_amd.make_params(4096, 128, 100, 0, 18)
    n = 10000
    C = 3
    assert cucm_amd.compress_bound(n, p) == 21 + 4 * 1 + 4 * C + C * (32 * 4096 + 1)


def test_peek_header_on_fixture():
    blob = open(os.path.join(os.path.dirname(__file__), "golden", "markov2_64k.cucm"), "rb").read()
    n, p = cucm_amd.peek_header(blob)
    assert n == 65536 and p.chunk_size == 16384 and p.unit_bits == 128 and p.table_bits == 18
    with pytest.raises(cucm_amd.CucmError):
        cucm_amd.peek_header(b"\x00" * 5)


def test_strerror():
    L = cucm_amd.lib()
    assert L.cucm_strerror(-3) == b"output buffer too small"
    assert L.cucm_strerror(0) == b"ok"


def test_null_arguments_are_rejected_without_device_work():
    L = cucm_amd.lib()
    assert L.cucm_ctx_create(None, 0, 0) == _native.E_PARAM
    assert L.cucm_ctx_set_max_waves(None, 4) == _native.E_PARAM
    assert L.cucm_ctx_last_stats(None, None) == _native.E_PARAM
    assert L.cucm_submit(None, None, None) == _native.E_PARAM
