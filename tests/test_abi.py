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


def test_c_struct_layouts_match_binding(tmp_path):
    """Every struct of include/lc_hip.h, compiled by the C compiler, has the size the binding uses."""
    pairs = {"lc_hip_ray": lc.RAY_DTYPE.itemsize, "lc_hip_surface_hit": lc.HIT_DTYPE.itemsize,
             "lc_hip_accel_modification": ctypes.sizeof(lc.Modification),
             "lc_hip_accel_option": ctypes.sizeof(lc.AccelOption), "lc_hip_ray_batch": ctypes.sizeof(lc.RayBatch),
             "lc_hip_bvh_stats": ctypes.sizeof(lc.BvhStats), "lc_hip_pt_instance": ctypes.sizeof(lc.PtInstance),
             "lc_hip_pt_scene": ctypes.sizeof(lc.PtScene), "lc_hip_pt_options": ctypes.sizeof(lc.PtOptions),
             "lc_hip_server_options": ctypes.sizeof(lc.ServerOptions),
             "lc_hip_render_request": ctypes.sizeof(lc.RenderRequest),
             "lc_hip_request_status": ctypes.sizeof(lc.RequestStatus),
             "lc_hip_serving_stats": ctypes.sizeof(lc.ServingStats),
             "lc_hip_trace_steps": ctypes.sizeof(lc.TraceSteps)}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "lc_hip.h"\nint main(void) {\n' +
                   "".join(f'    printf("{k} %zu\\n", sizeof({k}));\n' for k in pairs) + "    return 0;\n}\n")
    exe = tmp_path / "sizes"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {l.split()[0]: int(l.split()[1]) for l in out if l.strip()}
    assert got == pairs


def test_no_gpu_here_fails_loudly():
    if lc.lib().lc_hip_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lc.LCError):
        lc.Device(0)


def test_unknown_backend_name():
    with pytest.raises(lc.LCError):
        lc.create_device("cuda")
