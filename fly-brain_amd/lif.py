"""``LIFHip``: the reference's LIFWasm/LIFGpu model interface over the HIP engine.

Mirrors src/lifwasm.js:36-87 and src/lifgpu.js:183-370 (Lulzx/fly-brain):
the same constructor inputs (graph arrays, params merged over DEFAULTS, seed),
the same methods (step, setDrive, setDriveOne, setBias, setThr, addG, pulse,
reset, setBackground, setParams, flush, destroy, wake, nAwake) and the same
state fields (v, gE, gI, refr, trace, adapt, res, bias, thr, drive,
spikeCount, drivenSet, t, nslots, p).

The device holds the truth.  Reading a state field copies it from the device
(like LIFGpu's readback shadows); writing elements of the returned array does
not write back -- use the setters, as with LIFGpu.  ``step()`` returns the
spike list of the step in lif.c's order.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _l
from .params import merged, nslots as _nslots, to_struct


class HipGraph:
    """The shared read-only connectome on the device (refcounted)."""

    def __init__(self, N: int, indptr, indices, weights, sign, counts=None):
        lib = _l.load()
        self.N = int(N)
        self._indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
        self._indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self._W = np.ascontiguousarray(weights, dtype=np.float32)
        self._sign = np.ascontiguousarray(sign, dtype=np.float32)
        self._counts = None if counts is None else np.ascontiguousarray(counts, dtype=np.int32)
        self.E = int(self._indices.size)
        if self._indptr.size != self.N + 1 or self._W.size != self.E or self._sign.size != self.N:
            raise ValueError("graph array sizes do not match N / E")
        h = ctypes.c_void_p()
        _l.check(lib.fb_graph_create(self.N, self.E, _l.ptr(self._indptr), _l.ptr(self._indices), _l.ptr(self._W),
                                     _l.ptr(self._sign), _l.ptr(self._counts), ctypes.byref(h)))
        self.h = h

    @classmethod
    def from_synth(cls, g, W=None, sign=None, with_counts=True):
        return cls(g.N, g.indptr, g.indices, g.weights if W is None else W, g.sign if sign is None else sign,
                   g.counts if with_counts else None)

    def release(self):
        if getattr(self, "h", None):
            _l.check(_l.load().fb_graph_release(self.h))
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class LIFHip:
    """One brain on the GPU with the LIFWasm method set."""

    def __init__(self, graph: HipGraph, params: dict | None = None, seed=1, engine: str = "auto"):
        self.lib = _l.load()
        self.graph = graph
        self.N = graph.N
        self.p = merged(params)
        self.nslots = _nslots(self.p)
        self.engine = _l.ENGINES[engine]
        self.t = 0.0
        h = ctypes.c_void_p()
        ps = to_struct(self.p)
        _l.check(self.lib.fb_brain_create(graph.h, ctypes.byref(ps), float(seed), ctypes.byref(h)))
        self.h = h
        self.drivenSet: list[int] = []
        self._drive = np.zeros(self.N, dtype=np.float32)
        self._bias = np.zeros(self.N, dtype=np.float32)
        self._thr = np.zeros(self.N, dtype=np.float32)

    # ---------------------------------------------------------------- lifecycle
    def destroy(self):
        if getattr(self, "h", None):
            _l.check(self.lib.fb_brain_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def flush(self):
        _l.check(self.lib.fb_synchronize(self.h))

    def wake(self, *_):
        pass

    @property
    def nAwake(self) -> int:
        return self.N

    # ---------------------------------------------------------------- state fields
    def _get(self, which: str) -> np.ndarray:
        dt = np.uint32 if which == "spikeCount" else np.float32
        out = np.empty(self.N, dtype=dt)
        _l.check(self.lib.fb_get_array(self.h, _l.ARRAYS.index(which), _l.ptr(out)))
        return out

    def set_array(self, which: str, values):
        dt = np.uint32 if which == "spikeCount" else np.float32
        a = np.ascontiguousarray(values, dtype=dt)
        if a.size != self.N:
            raise ValueError("array size must be N")
        _l.check(self.lib.fb_set_array(self.h, _l.ARRAYS.index(which), _l.ptr(a)))

    def set_rng(self, state: int):
        """The xorshift32 state itself (a lane of the cord ensemble starts
        from a raw state rather than the seed hash)."""
        _l.check(self.lib.fb_set_rng(self.h, ctypes.c_uint32(int(state) & 0xFFFFFFFF)))

    def set_pool_feedback(self, pool_of, targets, base: float, gain: float, delay: int, record: int,
                          n_pools: int | None = None):
        """The closed pool loop (fb_set_pool_feedback): after every step the
        step's spikes are tallied by pool (pool_of[N], >= n_pools = none),
        recorded, and ``targets`` driven at base + gain * (the pools' total
        ``delay`` steps back) for the next step.  pool_of None turns it off."""
        if pool_of is None:
            _l.check(self.lib.fb_set_pool_feedback(self.h, None, 0, None, 0, 0.0, 0.0, 1, 0))
            self._n_pools = 0
            return
        pool = np.ascontiguousarray(pool_of, dtype=np.uint8)
        if pool.size != self.N:
            raise ValueError("pool_of must have N entries")
        self._n_pools = int(n_pools if n_pools is not None else pool[pool < 255].max() + 1)
        tg = np.ascontiguousarray(targets, dtype=np.int32)
        _l.check(self.lib.fb_set_pool_feedback(self.h, _l.ptr(pool), self._n_pools, _l.ptr(tg), int(tg.size),
                                               float(base), float(gain), int(delay), int(record)))

    def pool_counts(self) -> np.ndarray:
        """The recorded per-step pool counts, [steps][pools] uint16."""
        n = ctypes.c_int32()
        _l.check(self.lib.fb_pool_counts(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, self._n_pools), np.uint16)
        _l.check(self.lib.fb_pool_counts(self.h, _l.ptr(out), n.value, ctypes.byref(n)))
        return out

    v = property(lambda self: self._get("v"))
    gE = property(lambda self: self._get("gE"))
    gI = property(lambda self: self._get("gI"))
    refr = property(lambda self: self._get("refr"))
    trace = property(lambda self: self._get("trace"))
    adapt = property(lambda self: self._get("adapt"))
    res = property(lambda self: self._get("res"))
    spikeCount = property(lambda self: self._get("spikeCount"))
    # host-authoritative like LIFGpu (src/lifgpu.js:218-221)
    drive = property(lambda self: self._drive)
    bias = property(lambda self: self._bias)
    thr = property(lambda self: self._thr)

    @property
    def rng(self) -> int:
        r = ctypes.c_uint32()
        _l.check(self.lib.fb_brain_scalars(self.h, ctypes.byref(r), None, None, None))
        return int(r.value)

    @property
    def head(self) -> int:
        h = ctypes.c_int32()
        _l.check(self.lib.fb_brain_scalars(self.h, None, ctypes.byref(h), None, None))
        return int(h.value)

    def counters(self) -> dict:
        e, y, s, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _l.check(self.lib.fb_brain_counters(self.h, ctypes.byref(e), ctypes.byref(y), ctypes.byref(s),
                                            ctypes.byref(n)))
        return {"edge_events": e.value, "synapse_events": y.value, "spikes": s.value, "steps": n.value}

    def stat(self, name: str) -> float:
        v = ctypes.c_double()
        _l.check(self.lib.fb_brain_stat(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_stream(self, stream_handle: int | None):
        """Queue this brain's work on an external HIP stream (e.g. torch's)."""
        _l.check(self.lib.fb_brain_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    # ---------------------------------------------------------------- setters (src/lifwasm.js:69-76)
    def setDriveOne(self, i: int, rate: float):
        i = int(i)
        _l.check(self.lib.fb_set_drive_one(self.h, i, float(rate)))
        self._drive[i] = rate
        if rate > 0:
            if i not in self.drivenSet:
                self.drivenSet.append(i)
        elif i in self.drivenSet:
            self.drivenSet.remove(i)

    def setDrive(self, ix, rate: float):
        idx = np.ascontiguousarray(list(ix) if not isinstance(ix, np.ndarray) else ix, dtype=np.int32)
        _l.check(self.lib.fb_set_drive(self.h, _l.ptr(idx), int(idx.size), float(rate)))
        for i in idx.tolist():
            self._drive[i] = rate
            if rate > 0:
                if i not in self.drivenSet:
                    self.drivenSet.append(i)
            elif i in self.drivenSet:
                self.drivenSet.remove(i)

    def setBias(self, ix, mv: float):
        idx = np.ascontiguousarray(list(ix) if not isinstance(ix, np.ndarray) else ix, dtype=np.int32)
        _l.check(self.lib.fb_set_bias(self.h, _l.ptr(idx), int(idx.size), float(mv)))
        self._bias[idx] = mv

    def setThr(self, i: int, mv: float):
        _l.check(self.lib.fb_set_thr(self.h, int(i), float(mv)))
        self._thr[int(i)] = mv

    def addG(self, i: int, e: float, ii: float = 0.0):
        _l.check(self.lib.fb_add_g(self.h, int(i), float(e), float(ii or 0.0)))

    # fb_apply_deltas kinds (include/flybrain.h)
    DELTA_DRIVE, DELTA_BIAS, DELTA_THR, DELTA_GE, DELTA_GI = range(5)

    def apply_deltas(self, idx, kind, val):
        """A queue of setter calls in one launch: element q acts as setDriveOne,
        setBias (one index), setThr, or addG on one conductance, in order."""
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        kd = np.ascontiguousarray(kind, dtype=np.int32)
        vl = np.ascontiguousarray(val, dtype=np.float64)
        if not (ix.size == kd.size == vl.size):
            raise ValueError("idx, kind and val must have one entry per call")
        _l.check(self.lib.fb_apply_deltas(self.h, int(ix.size), _l.ptr(ix), _l.ptr(kd), _l.ptr(vl)))
        for i, k, v in zip(ix.tolist(), kd.tolist(), vl.tolist()):
            if k == self.DELTA_DRIVE:
                self._drive[i] = v
                if v > 0:
                    if i not in self.drivenSet:
                        self.drivenSet.append(i)
                elif i in self.drivenSet:
                    self.drivenSet.remove(i)
            elif k == self.DELTA_BIAS:
                self._bias[i] = v
            elif k == self.DELTA_THR:
                self._thr[i] = v

    def pulse(self, ix, mv: float):
        idx = np.ascontiguousarray(list(ix) if not isinstance(ix, np.ndarray) else ix, dtype=np.int32)
        _l.check(self.lib.fb_pulse(self.h, _l.ptr(idx), int(idx.size), float(mv)))

    def setBackground(self, rate_hz: float, amp_mv: float):
        self.p["bgRate"], self.p["bgAmp"] = rate_hz, amp_mv
        _l.check(self.lib.fb_set_background(self.h, float(rate_hz), float(amp_mv)))

    def setParams(self, q: dict):
        self.p.update(q)
        ps = to_struct(self.p)
        _l.check(self.lib.fb_set_params(self.h, ctypes.byref(ps)))
        self.nslots = _nslots(self.p)

    def reset(self):
        _l.check(self.lib.fb_reset(self.h))
        self.t = 0.0

    # ---------------------------------------------------------------- stepping
    def step(self) -> np.ndarray:
        nf = ctypes.c_int32()
        _l.check(self.lib.fb_step(self.h, 1, self.engine, ctypes.byref(nf)))
        self.t += self.p["dt"]
        return self.fired()

    def fired(self) -> np.ndarray:
        n = ctypes.c_int32()
        _l.check(self.lib.fb_fired(self.h, None, 0, ctypes.byref(n)))
        out = np.empty(max(1, n.value), dtype=np.int32)
        _l.check(self.lib.fb_fired(self.h, _l.ptr(out), int(out.size), ctypes.byref(n)))
        return out[: n.value]

    def advance(self, steps: int, engine: str | None = None) -> int:
        """Run ``steps`` steps without reading anything back; returns the
        spike count of the last step."""
        nf = ctypes.c_int32()
        eng = self.engine if engine is None else _l.ENGINES[engine]
        _l.check(self.lib.fb_step(self.h, int(steps), eng, ctypes.byref(nf)))
        self.t += steps * self.p["dt"]
        return int(nf.value)

    def run_async(self, steps: int, engine: str | None = None):
        """Queue ``steps`` steps and return at once (for timing)."""
        eng = self.engine if engine is None else _l.ENGINES[engine]
        _l.check(self.lib.fb_step(self.h, int(steps), eng, None))
        self.t += steps * self.p["dt"]

    def run_record(self, steps: int, engine: str | None = None, cap: int | None = None):
        """Per-step spike counts and the concatenated spike lists of ``steps`` steps."""
        eng = self.engine if engine is None else _l.ENGINES[engine]
        nf = np.zeros(max(1, steps), dtype=np.int32)
        cap = steps * 64 if cap is None else cap
        ras = np.zeros(max(1, cap), dtype=np.int32)
        tot = ctypes.c_int64()
        _l.check(self.lib.fb_run_record(self.h, int(steps), eng, _l.ptr(nf), _l.ptr(ras), int(ras.size),
                                        ctypes.byref(tot)))
        self.t += steps * self.p["dt"]
        if tot.value > ras.size:
            raise _l.FlyBrainError(4, f"raster capacity {ras.size} < {tot.value} spikes; pass cap=")
        return nf[:steps], ras[: tot.value]
