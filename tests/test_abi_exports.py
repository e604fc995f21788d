"""The C-ABI library loads without a GPU and exports every function that
include/naer/naer.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

from naer_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "naer", "naer.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(naer_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("naer_ctx_create", "naer_net_create", "naer_net_build", "naer_net_step", "naer_net_sync",
                 "naer_net_query", "naer_net_read_spikes", "naer_net_destroy", "naer_last_error",
                 "naer_pack_create", "naer_arena_create", "naer_arena_admit", "naer_arena_retire"):
        assert must in names


def test_library_exports_every_declared_function():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libnaer.so not built")
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    declared = set(declared_functions())
    bound = set(_abi.EXPORTS) | {"naer_last_error", "naer_version"}
    assert declared <= bound, sorted(declared - bound)


def test_version_string_without_device():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libnaer.so not built")
    assert b"gfx950" in _abi.lib().naer_version()
