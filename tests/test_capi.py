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


def test_multi_map_geometry_arguments_rejected_before_any_device_work():
    """pe_geom_*_multi check the map list (indices in range, offsets
    ascending from 0, finite positive half-widths) before any device work."""
    import numpy as np
    import pe_amd._native as native

    lib = native.lib()
    P = native.ptr
    states = np.zeros((3, 4))
    param = np.array([-1.0, 2.0, 3.0, 1.0, 0.5, 0.3, 0.05])
    obstacles = np.zeros((2, 3))
    out = np.empty(3, np.uint8)

    def backups(map_of, offsets, halves, n_maps=2):
        m = np.ascontiguousarray(map_of, np.int32)
        o = np.ascontiguousarray(offsets, np.int64)
        h = np.ascontiguousarray(halves, np.float64)
        return lib.pe_geom_backups_multi(P(states), len(states), P(m), n_maps, P(obstacles),
                                         P(o), P(h), P(param), P(out))

    for map_of, offsets, halves, text in [
            ([0, 1, 2], [0, 1, 2], [3.0, 3.0], b"map index"),
            ([0, -1, 1], [0, 1, 2], [3.0, 3.0], b"map index"),
            ([0, 1, 1], [1, 1, 2], [3.0, 3.0], b"start at 0"),
            ([0, 1, 1], [0, 2, 1], [3.0, 3.0], b"ascend"),
            ([0, 1, 1], [0, 1, 2], [3.0, 0.0], b"half-width"),
            ([0, 1, 1], [0, 1, 2], [3.0, np.nan], b"half-width")]:
        rc = backups(map_of, offsets, halves)
        assert rc == native.PE_ERR_INVALID, text
        assert text in lib.pe_last_error()
        with pytest.raises(ValueError):
            native.check(rc)
    controls = np.zeros((9, 2))
    ok = np.empty((3, 9), np.uint8)
    nxt = np.empty((3, 9, 4))
    m = np.array([0, 1, 1], np.int32)
    o = np.array([0, 1, 2], np.int64)
    h = np.array([3.0, 3.0])
    rc = lib.pe_geom_actions_multi(P(states), 3, P(m), 2, P(controls), 0, 0.05, P(obstacles),
                                   P(o), P(h), P(param), P(ok), P(nxt))
    assert rc == native.PE_ERR_INVALID and b"controls" in lib.pe_last_error()


def test_geometry_null_buffers_rejected_before_any_device_work():
    """Every geometry entry point refuses a null state or output buffer for
    n > 0 items with PE_ERR_INVALID, before it touches a device."""
    import numpy as np
    import pe_amd._native as native

    lib = native.lib()
    P = native.ptr
    null = None
    states = np.zeros((3, 4))
    param = np.array([-1.0, 2.0, 3.0, 1.0, 0.5, 0.3, 0.05])
    obstacles = np.zeros((2, 3))
    controls = np.zeros((9, 2))
    out_d, out_u8 = np.empty(3), np.empty(3, np.uint8)
    ok, nxt = np.empty((3, 9), np.uint8), np.empty((3, 9, 4))
    m = np.array([0, 1, 1], np.int32)
    o = np.array([0, 1, 2], np.int64)
    h = np.array([3.0, 3.0])
    calls = [
        lambda: lib.pe_geom_clearance(null, 3, P(obstacles), 2, 3.0, P(param), P(out_d)),
        lambda: lib.pe_geom_clearance(P(states), 3, P(obstacles), 2, 3.0, P(param), null),
        lambda: lib.pe_geom_backups(P(states), 3, P(obstacles), 2, 3.0, P(param), null),
        lambda: lib.pe_geom_actions(P(states), 3, P(controls), 9, 0.05, P(obstacles), 2, 3.0,
                                    P(param), null, P(nxt)),
        lambda: lib.pe_geom_actions(P(states), 3, null, 9, 0.05, P(obstacles), 2, 3.0,
                                    P(param), P(ok), P(nxt)),
        lambda: lib.pe_geom_connections(P(states), null, 3, P(obstacles), 2, 3.0, P(param),
                                        P(out_u8)),
        lambda: lib.pe_geom_backups_multi(null, 3, P(m), 2, P(obstacles), P(o), P(h), P(param),
                                          P(out_u8)),
        lambda: lib.pe_geom_actions_multi(P(states), 3, P(m), 2, P(controls), 9, 0.05,
                                          P(obstacles), P(o), P(h), P(param), P(ok), null),
    ]
    for k, call in enumerate(calls):
        assert call() == native.PE_ERR_INVALID, k
        assert b"null buffer" in lib.pe_last_error(), k


def test_hji5d_flag_bits_checked_before_any_device_work():
    """Unknown PE_HJI5D_* bits are rejected as invalid options; every known
    bit, PE_HJI5D_PSI_SINGLE included, passes the option check."""
    import pe_amd._native as native
    from pe_amd.reversible.free_dp import RelativeGrid, native_config
    from pe_amd.reversible.model import ReversibleConfig, VehicleLimits

    lib = native.lib()
    config = ReversibleConfig(name="a3", pursuer=VehicleLimits(-1.0, 2.0, 3.0, 1.0),
                              evader=VehicleLimits(-0.5, 1.0, 3.0, 2.0), capture_radius=0.3)
    known = (native.PE_HJI5D_LAYOUT_PLANE | native.PE_HJI5D_GENERIC |
             native.PE_HJI5D_NO_PREFETCH | native.PE_HJI5D_PSI_OUTER | native.PE_HJI5D_PSI_SINGLE)
    for flags, rejected in [(32, True), (known | 32, True), (native.PE_HJI5D_PSI_SINGLE, False),
                            (known, False)]:
        cfg = native_config(config, RelativeGrid(17, 12, 7, 7, 4.0))
        cfg.flags = flags
        h = ctypes.c_void_p()
        rc = lib.pe_hji5d_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc == 0:  # a device is present: the handle was made
            lib.pe_hji5d_destroy(h)
        invalid = rc == native.PE_ERR_INVALID and b"Invalid solver options" in lib.pe_last_error()
        assert invalid == rejected, (flags, rc, lib.pe_last_error())
