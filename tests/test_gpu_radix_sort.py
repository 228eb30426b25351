"""The library's own radix sort through lc_hip_radix_sort_pairs (lcub DeviceRadixSort::SortPairs,
include/luisa/backends/ext/cuda/lcub/device_radix_sort.h:15-27): the output equals numpy's stable
argsort of the keys (by the selected bits) applied to keys and values, for ragged sizes, bit
ranges, identity values, device-side counts below and above the capacity, and graph replays."""
import numpy as np
import pytest

import luisacompute_amd as lc

pytestmark = pytest.mark.gpu
DRS = lc.DeviceRadixSort


def _expect(keys, vals, begin, end):
    sel = (keys.astype(np.uint64) >> begin) & ((1 << (end - begin)) - 1)
    o = np.argsort(sel, kind="stable")
    return keys[o], (vals[o] if vals is not None else o.astype(np.uint32))


def _buffers(device, stream, keys, vals, cap):
    kin = device.create_buffer(4, cap)
    stream << kin.copy_from(keys)
    vin = None
    if vals is not None:
        vin = device.create_buffer(4, cap)
        stream << vin.copy_from(vals)
    kout, vout = device.create_buffer(4, cap), device.create_buffer(4, cap)
    temp = device.create_buffer(1, DRS.temp_storage_size(cap))
    return kin, vin, kout, vout, temp


def _read(stream, buf, n):
    out = np.zeros(n, np.uint32)
    stream << buf.copy_to(out) << lc.synchronize()
    return out


@pytest.mark.parametrize("n,begin,end,kind", [(1, 0, 32, "random"), (63, 0, 32, "random"), (4097, 0, 32, "random"),
                                              (1000003, 0, 32, "random"), (1000003, 5, 19, "random"),
                                              (1000003, 0, 30, "few"), (300000, 0, 24, "equal"),
                                              (300000, 0, 32, "sorted"), (300000, 0, 32, "reversed"),
                                              (8388608, 0, 24, "random")])
def test_sort_pairs_matches_stable_argsort(device, stream, n, begin, end, kind):
    print(f"[radix] n={n} bits=[{begin},{end}) keys={kind}", flush=True)
    rng = np.random.default_rng(n + begin)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if kind == "few":
        keys = rng.integers(0, 5, n, dtype=np.uint32) << 20
    elif kind == "equal":
        keys = np.full(n, 0x00ABCDEF, np.uint32)
    elif kind == "sorted":
        keys = np.sort(keys)
    elif kind == "reversed":
        keys = np.sort(keys)[::-1].copy()
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    kin, vin, kout, vout, temp = _buffers(device, stream, keys, vals, n)
    DRS.sort_pairs(stream, temp, kin, kout, vin, vout, n, begin, end)
    want_k, want_v = _expect(keys, vals, begin, end)
    np.testing.assert_array_equal(_read(stream, kout, n), want_k)
    np.testing.assert_array_equal(_read(stream, vout, n), want_v)
    np.testing.assert_array_equal(_read(stream, kin, n), keys)  # the input keys are left unchanged
    # identity values: the sorted permutation itself
    DRS.sort_pairs(stream, temp, kin, kout, None, vout, n, begin, end)
    np.testing.assert_array_equal(_read(stream, vout, n), _expect(keys, None, begin, end)[1])


def test_sort_reads_the_count_on_the_device(device, stream):
    """Only the first min(count, capacity) items are sorted; the outputs past them keep their words."""
    cap = 200000
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << 24, cap, dtype=np.uint32)
    kin, _, kout, vout, temp = _buffers(device, stream, keys, None, cap)
    cb = device.create_buffer(4, 2)
    for count in (0, 1, 4095, 4096, 4097, 123457, cap, cap + 999):
        print(f"[radix count] count={count}", flush=True)
        stream << cb.copy_from(np.array([0, count], np.uint32))
        stream << kout.copy_from(np.full(cap, 0xDEADBEEF, np.uint32)) << vout.copy_from(np.full(cap, 0xDEADBEEF, np.uint32))
        DRS.sort_pairs(stream, temp, kin, kout, None, vout, cap, 0, 24, count_buffer=cb, count_index=1)
        m = min(count, cap)
        want_k, want_v = _expect(keys[:m], None, 0, 24)
        gk, gv = _read(stream, kout, cap), _read(stream, vout, cap)
        np.testing.assert_array_equal(gk[:m], want_k)
        np.testing.assert_array_equal(gv[:m], want_v)
        assert (gk[m:] == 0xDEADBEEF).all() and (gv[m:] == 0xDEADBEEF).all()


def test_sort_replays_in_a_graph(device):
    """A captured sort replays with new keys and counts, eager sorts between the replays."""
    s = device.create_stream()
    cap = 100000
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1 << 32, cap, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(cap, dtype=np.uint32) * 3
    kin, vin, kout, vout, temp = _buffers(device, s, keys, vals, cap)
    cb = device.create_buffer(4, 1)
    s << cb.copy_from(np.array([cap], np.uint32))
    s.synchronize()
    with s.capture() as cap_:
        DRS.sort_pairs(s, temp, kin, kout, vin, vout, cap, 0, 32, count_buffer=cb)
    g = cap_.graph
    for k, count in enumerate((cap, 5000, 77777, cap)):
        print(f"[radix graph] replay {k} count={count}", flush=True)
        keys = rng.integers(0, 1 << 32, cap, dtype=np.uint64).astype(np.uint32)
        s << kin.copy_from(keys) << cb.copy_from(np.array([count], np.uint32))
        g.launch(s)
        want_k, want_v = _expect(keys[:count], vals[:count], 0, 32)
        np.testing.assert_array_equal(_read(s, kout, cap)[:count], want_k)
        np.testing.assert_array_equal(_read(s, vout, cap)[:count], want_v)
        # an eager sort on the same stream between replays
        DRS.sort_pairs(s, temp, kin, kout, None, vout, cap, 8, 16)
        np.testing.assert_array_equal(_read(s, vout, cap), _expect(keys, None, 8, 16)[1])
    del g, cap_


def test_sort_argument_errors(device, stream):
    n = 1000
    keys = np.arange(n, dtype=np.uint32)
    kin, vin, kout, vout, temp = _buffers(device, stream, keys, keys, n)
    for b, e in ((0, 0), (8, 4), (-1, 8), (0, 33)):
        with pytest.raises(lc.LCError, match="bits"):
            DRS.sort_pairs(stream, temp, kin, kout, vin, vout, n, b, e)
    with pytest.raises(lc.LCError, match="distinct"):
        DRS.sort_pairs(stream, temp, kin, kin, vin, vout, n)
    with pytest.raises(lc.LCError, match="fewer than"):
        DRS.sort_pairs(stream, temp, kin, kout, vin, vout, n + 1)
    small = device.create_buffer(1, 256)
    with pytest.raises(lc.LCError, match="temp buffer"):
        DRS.sort_pairs(stream, small, kin, kout, vin, vout, n)
    DRS.sort_pairs(stream, temp, kin, kout, vin, vout, 0)  # nothing to sort
