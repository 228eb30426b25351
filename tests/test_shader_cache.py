"""The shader code-object cache (hiprtc runs once per source; no GPU needed): a second compile in
the process is a memory hit, a new process finds the code object on disk, a different source or a
damaged cache file compiles again, LUISA_HIP_CACHE_DIR=0 keeps nothing on disk and
LUISA_DUMP_SOURCE=1 writes the source beside the code object."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SOURCE = r'''
#include "lc_hip_device.h"
struct P { alignas(16) lchip::dev::Buffer<float> out; alignas(16) uint4 ls_kid; };
extern "C" __global__ void k(const P p) {
    if (!lchip::dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = lchip::dev::dispatch_id().x;
    p.out[i] = 2.0f * (float)i + %s;
}
'''

SCRIPT = r'''
import json, sys
sys.path.insert(0, %r)
import luisacompute_amd as lc
src = sys.argv[1]
sizes = [lc.compile_shader_source(src) for _ in range(int(sys.argv[2]))]
h, m = lc.shader_cache_stats()
print(json.dumps({"hits": h, "misses": m, "sizes": sizes, "key": "%%016x" %% lc.shader_cache_key(src)}))
''' % ROOT


def _run(tmp_path, src, times, **env):
    e = dict(os.environ)
    e.update(env)
    out = subprocess.run([sys.executable, "-c", SCRIPT, src, str(times)], env=e, capture_output=True, text=True,
                         timeout=300, check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def test_memory_then_disk_hits(tmp_path):
    cache = tmp_path / "cache"
    src = SOURCE % "1.0f"
    r = _run(tmp_path, src, 2, LUISA_HIP_CACHE_DIR=str(cache))
    assert (r["misses"], r["hits"]) == (1, 1) and r["sizes"][0] == r["sizes"][1] > 1000
    path = cache / f"kernel_{r['key']}.hsaco"
    assert path.exists() and path.stat().st_size > r["sizes"][0]
    r2 = _run(tmp_path, src, 1, LUISA_HIP_CACHE_DIR=str(cache))
    assert (r2["misses"], r2["hits"]) == (0, 1) and r2["sizes"][0] == r["sizes"][0]
    # another source: another key, compiled
    r3 = _run(tmp_path, SOURCE % "3.0f", 1, LUISA_HIP_CACHE_DIR=str(cache))
    assert r3["key"] != r["key"] and (r3["misses"], r3["hits"]) == (1, 0)
    # a damaged file is ignored and rewritten
    data = path.read_bytes()
    path.write_bytes(data[:40])
    r4 = _run(tmp_path, src, 1, LUISA_HIP_CACHE_DIR=str(cache))
    assert (r4["misses"], r4["hits"]) == (1, 0)
    assert path.read_bytes() == data


def test_disk_cache_off_and_source_dump(tmp_path):
    src = SOURCE % "5.0f"
    r = _run(tmp_path, src, 1, LUISA_HIP_CACHE_DIR="0", HOME=str(tmp_path / "home"))
    assert r["misses"] == 1 and not (tmp_path / "home" / ".cache" / "luisacompute_amd").exists()
    cache = tmp_path / "dump"
    r = _run(tmp_path, src, 1, LUISA_HIP_CACHE_DIR=str(cache), LUISA_DUMP_SOURCE="1")
    assert (cache / f"kernel_{r['key']}.hip").read_text() == src
