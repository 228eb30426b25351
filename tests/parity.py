"""Parity helpers: compare backend hits with the oracle's under the stated tolerances.

Bar (from the workload definition):
  * instance and primitive ids bit-exact for rays whose two nearest hits differ by more than
    1e-6 relative in t (the "tie margin"; the oracle reports the second-nearest distance t2);
  * t and barycentrics within 1e-5 relative;
  * any-hit booleans bit-exact.
"""
import numpy as np

TIE_REL = 1e-6
VAL_REL = 1e-5
MISS = 0xFFFFFFFF


def compare_closest(gpu, ref2, rays=None):
    """gpu: HIT_DTYPE array; ref2: oracle HIT2_DTYPE array. Returns a report dict; raises on failure."""
    g_inst, g_prim = gpu["inst"], gpu["prim"]
    g_t = gpu["committed_ray_t"]
    g_u, g_v = gpu["bary"][:, 0], gpu["bary"][:, 1]
    r_hit = ref2["prim"] != MISS
    t, t2 = ref2["t"].astype(np.float64), ref2["t2"].astype(np.float64)
    tie = r_hit & np.isfinite(t2) & ((t2 - t) <= TIE_REL * np.abs(t))
    strict = ~tie
    n = len(gpu)
    # hit/miss and ids outside the tie margin
    id_ok = (g_inst == ref2["inst"]) & (g_prim == ref2["prim"])
    bad_ids = np.flatnonzero(strict & ~id_ok)
    # values: relative to the reference value (absolute floor 1e-30 for exact zeros)
    def rel(a, b):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
        return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    hit_strict = strict & r_hit
    t_err = rel(g_t, ref2["t"])
    bad_t = np.flatnonzero(strict & (t_err > VAL_REL))
    # barycentrics: relative error bounded against max(|ref|, 1e-3) so values near 0 do not
    # blow the relative error up (|u| <= 1 for hits)
    def bary_err(a, b):
        return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b.astype(np.float64)), 1e-3)
    bad_uv = np.flatnonzero(hit_strict & ((bary_err(g_u, ref2["u"]) > VAL_REL) | (bary_err(g_v, ref2["v"]) > VAL_REL)))
    # tie rays: the backend must still report a hit at the same distance
    bad_tie = np.flatnonzero(tie & ((g_prim == MISS) | (t_err > VAL_REL)))
    bit_exact = int(np.sum(hit_strict & (g_t == ref2["t"]) & (g_u == ref2["u"]) & (g_v == ref2["v"])))
    report = dict(rays=n, hits=int(r_hit.sum()), tie_excluded=int(tie.sum()), id_mismatch=len(bad_ids),
                  t_mismatch=len(bad_t), bary_mismatch=len(bad_uv), tie_mismatch=len(bad_tie),
                  bit_exact_hits=bit_exact, strict_hits=int(hit_strict.sum()))
    if len(bad_ids) or len(bad_t) or len(bad_uv) or len(bad_tie):
        i = int((list(bad_ids) + list(bad_t) + list(bad_uv) + list(bad_tie))[0])
        detail = f"first bad ray {i}: gpu {gpu[i]} oracle {ref2[i]}"
        if rays is not None:
            detail += f" ray {rays[i]}"
        raise AssertionError(f"closest-hit parity failed: {report}; {detail}")
    return report


def compare_any(gpu_any, ref_any):
    bad = np.flatnonzero(gpu_any.astype(np.uint8) != ref_any.astype(np.uint8))
    if len(bad):
        raise AssertionError(f"any-hit parity failed on {len(bad)} of {len(ref_any)} rays (first {bad[:8]})")
    return dict(rays=len(ref_any), occluded=int(ref_any.sum()))


def random_rigid(rng, lo=-50.0, hi=50.0, smin=0.5, smax=2.0):
    # uniform rotation from a uniform unit quaternion (Shoemake 1992)
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    s = rng.uniform(smin, smax)
    m = np.eye(4)
    m[:3, :3] = R * s
    m[:3, 3] = rng.uniform(lo, hi, 3)
    return m[:3, :4].astype(np.float32).reshape(12)


def random_boxes(rng, n, lo, hi, smin, smax):
    """(n, 6) float32 AABBs: centres uniform in [lo, hi]^3, per-axis extents uniform in [smin, smax]."""
    c = rng.uniform(lo, hi, (n, 3))
    e = rng.uniform(smin, smax, (n, 3)) * 0.5
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32)


def compare_query(gpu, ref, scene=None, rays=None, **policy):
    """gpu: COMMITTED_HIT_DTYPE array of the nearest-commit query; ref: the oracle's COMMITTED_DTYPE
    answer. Hit kind, ids, t and barycentrics bit-exact; where the ids differ both answers must be
    committed candidates (checked with the oracle when `scene` is given) at the same t within the
    tie margin. Returns a report dict; raises on failure."""
    g_t, r_t = gpu["committed_ray_t"], ref["t"]
    same_id = (gpu["inst"] == ref["inst"]) & (gpu["prim"] == ref["prim"]) & (gpu["hit_type"] == ref["hit_type"])
    bits = lambda a: np.ascontiguousarray(a, np.float32).view(np.uint32)
    exact = same_id & (bits(g_t) == bits(r_t)) & (bits(gpu["bary"][:, 0]) == bits(ref["u"])) & \
        (bits(gpu["bary"][:, 1]) == bits(ref["v"]))
    bad = []
    ties = 0
    for i in np.flatnonzero(~exact):
        if same_id[i]:
            bad.append(i)  # same candidate, different numbers
            continue
        both_hit = gpu["hit_type"][i] != 0 and ref["hit_type"][i] != 0
        tie = both_hit and abs(float(g_t[i]) - float(r_t[i])) <= TIE_REL * abs(float(r_t[i]))
        if tie and scene is not None:
            c = scene.query_candidate(rays[i], gpu["inst"][i], gpu["prim"][i], gpu["hit_type"][i], **policy)
            tie = c is not None and np.float32(c[0]) == g_t[i]
        if tie:
            ties += 1
        else:
            bad.append(i)
    if bad:
        i = bad[0]
        raise AssertionError(f"query parity failed on {len(bad)} of {len(gpu)} rays; first {i}: gpu {gpu[i]} "
                             f"oracle {ref[i]}")
    kinds = np.bincount(ref["hit_type"].astype(np.int64), minlength=3)
    return dict(rays=len(gpu), miss=int(kinds[0]), surface=int(kinds[1]), procedural=int(kinds[2]), ties=ties)
