"""The C-ABI library loads and exports every entry point declared in include/lc_hip.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

import luisacompute_amd as lc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lc_hip.h")).read()
    return sorted(set(re.findall(r"\b(lc_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "lc_hip_accel_trace_closest" in syms and "lc_hip_mesh_build" in syms
    assert len(syms) >= 40


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(lc.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, f"missing exports: {missing}"


def test_binding_covers_header():
    assert set(declared_symbols()) == set(lc.EXPORTED_SYMBOLS)


def test_abi_struct_sizes():
    assert lc.RAY_DTYPE.itemsize == 32          # include/luisa/runtime/rtx/ray.h:16
    assert lc.HIT_DTYPE.itemsize == 24          # include/luisa/runtime/rtx/hit.h:36
    assert ctypes.sizeof(lc.Modification) == 80  # command.h:534-551 (alignas 16)
    assert lc.HIT_DTYPE.fields["committed_ray_t"][1] == 16


def test_no_gpu_here_fails_loudly():
    if lc.lib().lc_hip_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lc.LCError):
        lc.Device(0)


def test_unknown_backend_name():
    with pytest.raises(lc.LCError):
        lc.create_device("cuda")
