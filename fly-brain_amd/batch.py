"""``LIFBatch``: B ensemble members of one brain model on the GPU.

The reference runs ensembles as one brain per worker over a shared graph
(src/brainsetup.js:8-27, scripts/lif_ensemble.mjs:16-35, scripts/run_ensemble.mjs)
and as a lane ensemble in Bend (kernel.bend:293-378, cordens.bend:1-60).  Here
the B members advance together: member b is the brain ``LIFHip`` would make
with seed ``seeds[b]`` and weights ``W = f32(count * gains[b, type(pre)])``
(writeGraph's outScale rule, src/lifwasm.js:23,31); members share the driven
set; each may have its own silenced neurons (``thr = 1e6``) and run length.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _l
from .params import merged, nslots as _nslots, to_struct


class LIFBatch:
    def __init__(self, graph, params: dict | None, seeds, type_of, gains):
        self.lib = _l.load()
        self.graph = graph
        self.N = graph.N
        self.p = merged(params)
        self.nslots = _nslots(self.p)
        seeds = np.ascontiguousarray(seeds, dtype=np.float64)
        self.B = int(seeds.size)
        self._type = np.ascontiguousarray(type_of, dtype=np.int32)
        g = np.ascontiguousarray(gains, dtype=np.float32).reshape(self.B, -1)
        self.nTypes = int(g.shape[1])
        self._gains = g
        h = ctypes.c_void_p()
        ps = to_struct(self.p)
        _l.check(self.lib.fb_batch_create(graph.h, ctypes.byref(ps), self.B, _l.ptr(seeds), _l.ptr(self._type),
                                          self.nTypes, _l.ptr(g), ctypes.byref(h)))
        self.h = h
        self.drivenSet: list[int] = []

    def destroy(self):
        if getattr(self, "h", None):
            _l.check(self.lib.fb_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def setDrive(self, ix, rate: float):
        idx = np.ascontiguousarray(ix, dtype=np.int32)
        _l.check(self.lib.fb_batch_set_drive(self.h, _l.ptr(idx), int(idx.size), float(rate)))
        for i in idx.tolist():
            if rate > 0:
                if i not in self.drivenSet:
                    self.drivenSet.append(i)
            elif i in self.drivenSet:
                self.drivenSet.remove(i)

    def setThr(self, member: int, ix, mv: float):
        idx = np.ascontiguousarray(ix, dtype=np.int32)
        _l.check(self.lib.fb_batch_set_member(self.h, 8, int(member), _l.ptr(idx), int(idx.size), float(mv)))

    def setBias(self, member: int, ix, mv: float):
        idx = np.ascontiguousarray(ix, dtype=np.int32)
        _l.check(self.lib.fb_batch_set_member(self.h, 7, int(member), _l.ptr(idx), int(idx.size), float(mv)))

    def silence(self, member: int, ix):
        """setThr(i, 1e6) on every listed neuron of one member (scripts/lif_ensemble.mjs:35)."""
        self.setThr(member, ix, 1e6)

    def set_lanes(self, sign_gain: bool, lane_drive=None):
        """Lane mode (fb_batch_set_lanes): the type gain scales the sign, and
        lane_drive [B][N] gives each member its own rates."""
        ld = None if lane_drive is None else np.ascontiguousarray(lane_drive, dtype=np.float32).reshape(self.B, -1)
        _l.check(self.lib.fb_batch_set_lanes(self.h, 1 if sign_gain else 0, _l.ptr(ld)))

    def set_shape(self, neurons_per_block: int):
        """The lookahead engine's regular-block shape (72, 80, 96 or 128 neurons per
        workgroup), or 0 for the autotuner's choice (fb_batch_set_shape)."""
        _l.check(self.lib.fb_batch_set_shape(self.h, int(neurons_per_block)))

    def set_rng(self, states):
        st = np.ascontiguousarray(states, dtype=np.uint32)
        _l.check(self.lib.fb_batch_set_rng(self.h, _l.ptr(st)))

    def set_pool_feedback(self, pool_of, n_pools: int, targets, base: float, gain: float, delays, record: int):
        # pool_of None clears the loop
        pool = None if pool_of is None else np.ascontiguousarray(pool_of, dtype=np.uint8)
        tg = np.ascontiguousarray(targets, dtype=np.int32)
        dl = np.ascontiguousarray(delays, dtype=np.int32)
        self._n_pools = int(n_pools)
        _l.check(self.lib.fb_batch_set_pool_feedback(self.h, _l.ptr(pool), int(n_pools), _l.ptr(tg), int(tg.size),
                                                     float(base), float(gain), _l.ptr(dl), int(record)))

    def pool_counts(self) -> np.ndarray:
        """[B][steps][pools] uint16 of the pool loop's record."""
        n = ctypes.c_int32()
        _l.check(self.lib.fb_batch_pool_counts(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros((self.B, n.value, self._n_pools), np.uint16)
        _l.check(self.lib.fb_batch_pool_counts(self.h, _l.ptr(out), n.value, ctypes.byref(n)))
        return out

    def set_steps(self, steps):
        st = np.ascontiguousarray(steps, dtype=np.int32)
        if st.size != self.B:
            raise ValueError("one step count per member")
        _l.check(self.lib.fb_batch_set_steps(self.h, _l.ptr(st)))

    def set_stream(self, handle: int | None):
        _l.check(self.lib.fb_batch_set_stream(self.h, ctypes.c_void_p(handle or 0)))

    def set_engine(self, engine: str):
        """'auto', 'launch' (per-step kernels) or 'persistent' (lookahead)."""
        _l.check(self.lib.fb_batch_set_engine(self.h, _l.ENGINES[engine]))

    def stat(self, name: str) -> float:
        v = ctypes.c_double()
        _l.check(self.lib.fb_batch_stat(self.h, name.encode(), ctypes.byref(v)))
        return float(v.value)

    def run(self, steps: int):
        _l.check(self.lib.fb_batch_run(self.h, int(steps), None, None, None, None, 0, None))

    def run_record(self, steps: int, cap: int | None = None):
        """Per step: the neurons that spiked in any member, their spike and forced masks."""
        cap = steps * 256 if cap is None else cap
        nact = np.zeros(max(1, steps), dtype=np.int32)
        ids = np.zeros(max(1, cap), dtype=np.int32)
        masks = np.zeros(max(1, cap), dtype=np.uint64)
        fm = np.zeros(max(1, cap), dtype=np.uint64)
        tot = ctypes.c_int64()
        _l.check(self.lib.fb_batch_run(self.h, int(steps), _l.ptr(nact), _l.ptr(ids), _l.ptr(masks), _l.ptr(fm),
                                       int(cap), ctypes.byref(tot)))
        if tot.value > cap:
            raise _l.FlyBrainError(4, f"record capacity {cap} < {tot.value}")
        return nact[:steps], ids[: tot.value], masks[: tot.value], fm[: tot.value]

    def member_rasters(self, nact, ids, masks, fmasks):
        """Each member's per-step spike lists in lif.c order: forced spikes in
        driven-set order, then threshold spikes ascending (src/wasm/lif.c:49-76)."""
        pos = {i: k for k, i in enumerate(self.drivenSet)}
        start = np.concatenate([[0], np.cumsum(nact)])
        out_nf = np.zeros((self.B, nact.size), dtype=np.int32)
        out_ras = [[] for _ in range(self.B)]
        for s in range(nact.size):
            sl = slice(start[s], start[s + 1])
            ii, mm, ff = ids[sl], masks[sl], fmasks[sl]
            for b in range(self.B):
                bit = np.uint64(1) << np.uint64(b)
                on = (mm & bit) != 0
                if not on.any():
                    continue
                forced = on & ((ff & bit) != 0)
                f_ids = sorted(ii[forced].tolist(), key=lambda i: pos[i])
                t_ids = sorted(ii[on & ~forced].tolist())
                lst = f_ids + t_ids
                out_nf[b, s] = len(lst)
                out_ras[b].extend(lst)
        return out_nf, [np.asarray(r, dtype=np.int32) for r in out_ras]

    def array(self, which: str, member: int) -> np.ndarray:
        names = ("v", "gE", "gI", "refr", "trace", "adapt", "res", "bias", "thr", "drive", "spikeCount")
        dt = np.uint32 if which == "spikeCount" else np.float32
        out = np.empty(self.N, dtype=dt)
        _l.check(self.lib.fb_batch_get_array(self.h, names.index(which), int(member), _l.ptr(out)))
        return out

    def rng(self) -> np.ndarray:
        out = np.empty(self.B, dtype=np.uint32)
        _l.check(self.lib.fb_batch_rng(self.h, _l.ptr(out)))
        return out

    def counters(self) -> dict:
        e, y, s, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _l.check(self.lib.fb_batch_counters(self.h, ctypes.byref(e), ctypes.byref(y), ctypes.byref(s), ctypes.byref(n)))
        return {"edge_events": e.value, "synapse_events": y.value, "spikes": s.value, "steps": n.value}
