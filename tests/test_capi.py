"""The C-ABI library loads and exports every symbol include/pe_amd.h declares.

No compute calls here (no GPU needed): only pe_abi_version / pe_last_error,
which never touch a device.
"""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(pe_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_functions():
    names = declared_functions()
    assert "pe_hji5d_create" in names and "pe_last_error" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    import pe_amd._native as native

    lib = native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_declares_every_header_function():
    import pe_amd._native as native

    assert set(declared_functions()) <= set(native.SIGNATURES)


def test_abi_version_and_error_text_without_device():
    import pe_amd._native as native

    lib = native.lib()
    assert lib.pe_abi_version() == 1
    assert isinstance(lib.pe_last_error(), bytes)


def test_invalid_options_rejected_before_any_device_work():
    """pe_hji5d_create validates its options first (native_relative.cpp:260-265)."""
    import pe_amd._native as native

    lib = native.lib()
    cfg = native.HJI5DConfig()
    cfg.positions = 2  # < 3
    h = ctypes.c_void_p()
    rc = lib.pe_hji5d_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == native.PE_ERR_INVALID
    assert b"Invalid solver options" in lib.pe_last_error()
    with pytest.raises(ValueError):
        native.check(rc)


def test_struct_mirrors_match_the_c_layout():
    import pe_amd._native as native

    lib = native.lib()
    for name, mirror in native.STRUCTS.items():
        assert lib.pe_abi_struct_size(name.encode()) == ctypes.sizeof(mirror), name
