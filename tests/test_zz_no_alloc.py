"""fb_apply_deltas_noalloc, the allocation-free form of fb_apply_deltas,
called through the C ABI: on two calls in a row (the second with a smaller
queue) the brain it acts on steps exactly as a brain given the same queues by
fb_apply_deltas -- spike lists, state arrays and the stream, bit for bit --
and a queue past its bound is refused before any device work."""
import ctypes
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lif = importlib.import_module("fly-brain_amd.lif")
L = importlib.import_module("fly-brain_amd.lib")
synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

KINDS = 5     # drive, bias, thr, gE, gI


def queue(rng, N, n):
    """A queue with repeats on a few neurons, every kind, drives switched on
    and off, and negative conductance additions."""
    idx = np.concatenate([rng.integers(0, N, n - 8), np.full(8, 7)]).astype(np.int32)
    kind = rng.integers(0, KINDS, n).astype(np.int32)
    val = np.where(kind == 0, rng.choice([0.0, 150.0, 400.0], n),
                   np.where(kind == 1, rng.uniform(0, 8, n),
                            np.where(kind == 2, rng.uniform(-2, 3, n), rng.uniform(-0.5, 1.5, n))))
    return idx, kind, val.astype(np.float64)


def call(lib, name, b, idx, kind, val):
    return getattr(lib, name)(b.h, int(idx.size), L.ptr(idx), L.ptr(kind), L.ptr(val))


def test_noalloc_deltas_equal_the_allocating_entry_point():
    print("fb_apply_deltas_noalloc vs fb_apply_deltas: two calls, the second smaller", flush=True)
    lib = L.load()
    g = synth.tiny_graph(3000, 30, seed=9, p_count=0.3)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.6).astype(np.float32)
    p = P.config_params()
    hg = lif.HipGraph(g.N, g.indptr, g.indices, g.weights, sign, g.counts)
    ref = lif.LIFHip(hg, p, seed=4, engine="launch")
    new = lif.LIFHip(hg, p, seed=4, engine="launch")
    rng = np.random.default_rng(11)
    for n in (5000, 700):
        idx, kind, val = queue(rng, g.N, n)
        assert call(lib, "fb_apply_deltas", ref, idx, kind, val) == 0
        assert call(lib, "fb_apply_deltas_noalloc", new, idx, kind, val) == 0, lib.fb_last_error()
        nr, rr = ref.run_record(150, cap=150 * g.N)
        nn, rn = new.run_record(150, cap=150 * g.N)
        assert np.array_equal(nr, nn) and np.array_equal(rr, rn)
        assert int(nr.sum()) > 0
        for k in ("v", "gE", "gI", "refr", "drive", "bias", "thr"):
            assert np.array_equal(ref._get(k).view(np.uint32), new._get(k).view(np.uint32)), (n, k)
        assert ref.rng == new.rng
        drv_r = np.zeros(g.N, np.int32)
        drv_n = np.zeros(g.N, np.int32)
        cr, cn = ctypes.c_int32(), ctypes.c_int32()
        assert lib.fb_driven(ref.h, L.ptr(drv_r), g.N, ctypes.byref(cr)) == 0
        assert lib.fb_driven(new.h, L.ptr(drv_n), g.N, ctypes.byref(cn)) == 0
        assert cr.value == cn.value and np.array_equal(drv_r[:cr.value], drv_n[:cn.value])
    # past the bound: refused, nothing applied
    big = np.zeros(65537, np.int32)
    assert call(lib, "fb_apply_deltas_noalloc", new, big, big, big.astype(np.float64)) != 0
    assert b"exceed" in lib.fb_last_error()
    ref.destroy()
    new.destroy()
    hg.release()
