"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle and
cross-checked against a double-precision brute force when written).

CPU: the oracle still reproduces every fixture bit for bit.
GPU: the HIP path reproduces every fixture (ids outside the tie margin, t/u/v within 1e-5)."""
import glob
import os

import numpy as np
import pytest

import luisacompute_amd as lc
from oracle import binding as ob
from parity import compare_any, compare_closest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    f = np.load(path, allow_pickle=False)
    return {k: f[k] for k in f.files}


def expected_hits(f):
    exp = np.zeros(len(f["rays"]), ob.HIT2_DTYPE)
    for k in ("inst", "prim", "u", "v", "t", "t2"):
        exp[k] = f[k]
    return exp


def test_fixtures_present():
    names = {os.path.basename(p) for p in FIXTURES}
    assert {"icosphere_l3_4096.npz", "soup_2000_4096.npz", "rtx_two_instances_64x64.npz"} <= names


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_fixture(path):
    f = load(path)
    m = ob.Mesh(f["verts"], f["tris"])
    sc = ob.Scene()
    for a, mk in zip(f["affines"], f["masks"]):
        sc.add(m, a, int(mk))
    sc.commit()
    h = sc.trace_closest2(f["rays"], mask=int(f["ray_mask"]))
    for k in ("inst", "prim", "u", "v", "t", "t2"):
        assert np.array_equal(h[k], f[k]), k
    assert np.array_equal(sc.trace_any(f["rays"], mask=int(f["ray_mask"])), f["any"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_gpu_reproduces_fixture(device, stream, path):
    from luisacompute_amd import host
    f = load(path)
    print(f"[fixture {os.path.basename(path)}]", flush=True)
    mesh = host.mesh_from_arrays(device, stream, f["verts"], f["tris"])
    accel = device.create_accel()
    for a, mk in zip(f["affines"], f["masks"]):
        accel.emplace_back(mesh, a, visibility_mask=int(mk))
    stream << accel.build()
    rays = host.as_rays(f["rays"])
    g = host.trace_closest(device, stream, accel, rays, vis_mask=int(f["ray_mask"]))
    rep = compare_closest(g, expected_hits(f), rays)
    ga = host.trace_any(device, stream, accel, rays, vis_mask=int(f["ray_mask"]))
    compare_any(ga, f["any"])
    assert rep["bit_exact_hits"] == rep["strict_hits"]
