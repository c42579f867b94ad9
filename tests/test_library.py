"""The C-ABI library (no GPU needed): it loads and exports every entry point
include/adgpu.h declares; the oracle library exports what its header declares."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(adgpu_\w+|orc_\w+)\s*\(", txt, flags=re.M)))


def test_header_symbols_exported():
    from adgpu import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.fail("libadgpu_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = _declared(os.path.join(REPO, "include", "adgpu.h"))
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_abi.EXPORTED)
    lib.adgpu_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.adgpu_version()


def test_null_arguments_are_rejected_without_gpu():
    from adgpu import _abi
    lib = _abi.load_library()
    assert lib.adgpu_open(0, None) == 1                      # ADGPU_ERR_ARG
    assert lib.adgpu_close(None) == 1
    assert lib.adgpu_job_destroy(None) == 1
    assert lib.adgpu_last_error(None) == b"null context"
    # serving loop and team table: argument checks come before any HIP call
    cfg = _abi.ServerConfig(4, 0, 0)
    h = ctypes.c_void_p()
    assert lib.adgpu_server_create(None, ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert lib.adgpu_server_submit(None, None, None) == 1
    assert lib.adgpu_server_next_done(None, 0, None) == 1
    assert lib.adgpu_server_result(None, 1, 0, None, None, None, None, None, None) == 1
    assert lib.adgpu_server_stats_get(None, None) == 1
    assert lib.adgpu_server_destroy(None) == 1
    assert lib.adgpu_set_team_table(None, 0, None, None) == 1
    assert lib.adgpu_set_ls_handoff(None, 32) == 1


def test_oracle_symbols():
    import oracle
    lib = oracle.lib()
    names = _declared(os.path.join(REPO, "oracle", "adock_oracle.h"))
    assert not [n for n in names if not hasattr(lib, n)]


def test_autotune_table_from_rates():
    """Classes bounded by the measured sizes; equal neighbours merge."""
    from adgpu.autotune import table_from
    rates = {24: {16: 1.0, 32: 3.0, 64: 2.0}, 32: {16: 1.0, 32: 3.0, 64: 2.0}, 40: {32: 1.0, 64: 2.0},
             64: {32: 1.0, 64: 2.0}, 96: {32: 5.0, 64: 2.0}}
    assert table_from(rates) == [(32, 32), (64, 64), (96, 32)]
    assert table_from({40: {32: 2.0, 64: 1.0}}) == [(40, 32)]
