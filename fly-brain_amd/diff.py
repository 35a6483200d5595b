"""``LIFDiffHip``: the reference's differentiable LIF (LIFDiff) on the GPU.

Mirrors ``LIFDiff`` of src/lifdiff.js:84-686: construct over a CSR graph with
the model options, ``setDrive`` / ``setDriveOne`` in insertion order, then
``forward(steps, seed)`` records a tape and ``backward(dLdSpike, truncate,
lossFrom)`` returns ``{params, logGain, edge, inScale, clamped}`` with the nine
scalar gradients in ``DIFF_PARAMS`` order.  ``edge`` (dL/d edgeGain per edge,
caller's edge order) and ``set_readout`` (a filtered-rate loss) have no
counterpart in the reference; ``type_grads`` turns the log-gain gradient into
type-gain gradients.

The tape lives on the device inside the object: ``forward`` returns the loss,
and ``backward`` differentiates the last recorded forward pass.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _l
from .params import DIFF_PARAMS, diff_merged, to_diff_struct

INPUTS = ("outScale", "logGain", "thrOffset", "sizeLog", "drive", "inScale", "surrogateBeta")
STATE = ("v", "gE", "gI", "adapt", "res", "refr", "spikeCount", "inScale")


def train_csr(train):
    """A spike train (a sequence over steps of index lists, or None) as
    (ptr int64 [steps + 1], idx int32); None gives an empty train."""
    if train is None:
        return np.zeros(1, np.int64), np.zeros(0, np.int32)
    lists = [np.asarray(x if x is not None else [], dtype=np.int64).ravel() for x in train]
    ptr = np.zeros(len(lists) + 1, np.int64)
    if lists:
        ptr[1:] = np.cumsum([x.size for x in lists])
    idx = np.concatenate(lists).astype(np.int32) if lists and ptr[-1] else np.zeros(0, np.int32)
    return ptr, idx

class LIFDiffHip:
    def __init__(self, N, indptr, indices, counts, sign, opts: dict | None = None):
        self.lib = _l.load()
        self.N = int(N)
        self.p = diff_merged(opts)
        self._indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
        self._indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.E = int(self._indices.size)
        c = np.ascontiguousarray(counts, dtype=np.float32)
        s = np.ascontiguousarray(sign, dtype=np.float32)
        if c.size != self.E or s.size != self.N or self._indptr.size != self.N + 1:
            raise ValueError("LIFDiffHip: array sizes do not match N and E")
        h = ctypes.c_void_p()
        ps = to_diff_struct(self.p)
        _l.check(self.lib.fb_diff_create(self.N, self.E, _l.ptr(self._indptr), _l.ptr(self._indices), _l.ptr(c),
                                         _l.ptr(s), ctypes.byref(ps), ctypes.byref(h)))
        self.h = h
        self.drivenSet: list[int] = []
        self.drive = np.zeros(self.N, np.float32)
        self.R = 0
        self.steps = 0

    def destroy(self):
        if getattr(self, "h", None):
            _l.check(self.lib.fb_diff_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ---------------------------------------------------------------- inputs
    def set_params(self, **over):
        self.p.update(over)
        ps = to_diff_struct(self.p)
        _l.check(self.lib.fb_diff_set_params(self.h, ctypes.byref(ps)))

    def set_input(self, name: str, values):
        """A per-neuron input.  ``inScale`` given is taken verbatim instead of
        the one formed from sizeLog; ``surrogateBeta`` is a per-neuron beta;
        None clears either (the derived inScale, the scalar beta)."""
        if values is None:
            if name not in ("inScale", "surrogateBeta"):
                raise ValueError(f"{name}: only inScale and surrogateBeta can be cleared")
            _l.check(self.lib.fb_diff_set_input(self.h, INPUTS.index(name), None))
            return
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.size != self.N:
            raise ValueError(f"{name}: expected {self.N} values")
        _l.check(self.lib.fb_diff_set_input(self.h, INPUTS.index(name), _l.ptr(a)))
        if name == "drive":
            self.drive = a.copy()

    def set_edge_gain(self, eg):
        if eg is None:
            _l.check(self.lib.fb_diff_set_edge_gain(self.h, None))
            return
        a = np.ascontiguousarray(eg, dtype=np.float32)
        if a.size != self.E:
            raise ValueError(f"edgeGain: expected {self.E} values")
        _l.check(self.lib.fb_diff_set_edge_gain(self.h, _l.ptr(a)))

    def set_masks(self, thr_mask=None, bias_mask=None):
        t = None if thr_mask is None else np.ascontiguousarray(thr_mask, dtype=np.uint8)
        b = None if bias_mask is None else np.ascontiguousarray(bias_mask, dtype=np.uint8)
        _l.check(self.lib.fb_diff_set_masks(self.h, None if t is None else _l.ptr(t), None if b is None else _l.ptr(b)))

    def set_driven(self, ix):
        a = np.ascontiguousarray(ix, dtype=np.int32)
        _l.check(self.lib.fb_diff_set_driven(self.h, _l.ptr(a), int(a.size)))
        self.drivenSet = a.tolist()

    def setDriveOne(self, i: int, rate: float):
        """setDriveOne (src/lifdiff.js:207): the rate, and drivenSet membership
        in insertion order."""
        self.drive[int(i)] = rate
        if rate > 0:
            if int(i) not in self.drivenSet:
                self.drivenSet.append(int(i))
        elif int(i) in self.drivenSet:
            self.drivenSet.remove(int(i))
        self.set_input("drive", self.drive)
        self.set_driven(self.drivenSet)

    def setDrive(self, ix, rate: float):
        for i in ix:
            self.drive[int(i)] = rate
            if rate > 0:
                if int(i) not in self.drivenSet:
                    self.drivenSet.append(int(i))
            elif int(i) in self.drivenSet:
                self.drivenSet.remove(int(i))
        self.set_input("drive", self.drive)
        self.set_driven(self.drivenSet)

    def set_loss(self, dLdSpike=None, loss_from: int = 0):
        a = None if dLdSpike is None else np.ascontiguousarray(dLdSpike, dtype=np.float32)
        _l.check(self.lib.fb_diff_set_loss(self.h, None if a is None else _l.ptr(a), int(loss_from)))

    def set_readout(self, idx, target_hz, tau_ms: float):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        t = np.ascontiguousarray(target_hz, dtype=np.float32)
        self.R = int(i.size)
        _l.check(self.lib.fb_diff_set_readout(self.h, _l.ptr(i), self.R, _l.ptr(t), float(tau_ms)))

    def set_drive_schedule(self, idx, epoch_steps: int, sched, drive_soft: bool = False):
        """driveIdx / driveEpoch (src/lifdiff.js:135-144): the rates sched[e][k]
        that the onEpoch hook gives neuron idx[k] at the start of epoch e,
        applied with setDriveOne's semantics; drive_soft = driveSoft.  None clears."""
        if idx is None:
            _l.check(self.lib.fb_diff_set_drive_schedule(self.h, None, 0, 1, None, 0, 1 if drive_soft else 0))
            self._dsched = None
            return
        i = np.ascontiguousarray(idx, dtype=np.int32)
        sc = np.ascontiguousarray(sched, dtype=np.float32).reshape(-1, i.size)
        self._dsched = sc.shape
        _l.check(self.lib.fb_diff_set_drive_schedule(self.h, _l.ptr(i), int(i.size), int(epoch_steps), _l.ptr(sc),
                                                      int(sc.shape[0]), 1 if drive_soft else 0))

    def set_loss_per_epoch(self, dper, epoch_steps: int):
        """dLdSpikePerEpoch (src/lifdiff.js:519-522): dper[e][i]; None clears."""
        if dper is None:
            _l.check(self.lib.fb_diff_set_loss_per_epoch(self.h, None, 0, int(epoch_steps)))
            return
        a = np.ascontiguousarray(dper, dtype=np.float32).reshape(-1, self.N)
        _l.check(self.lib.fb_diff_set_loss_per_epoch(self.h, _l.ptr(a), int(a.shape[0]), int(epoch_steps)))

    def drive_grad(self) -> np.ndarray:
        """dL/d sched [epochs][nD] of the last backward (LIFDiff.backward's `drive`)."""
        out = np.zeros(self._dsched, np.float32)
        _l.check(self.lib.fb_diff_drive_grad(self.h, _l.ptr(out)))
        return out

    def set_tape_capacity(self, per_step: int):
        _l.check(self.lib.fb_diff_set_tape_capacity(self.h, int(per_step)))

    def set_stream(self, handle: int | None):
        _l.check(self.lib.fb_diff_set_stream(self.h, ctypes.c_void_p(handle) if handle else None))

    # ---------------------------------------------------------------- passes
    def set_spike_train(self, train):
        """spikeTrain (src/lifdiff.js:298-307): train[t] lists the neurons forced
        to spike at step t (None or empty for none); None clears."""
        ptr, idx = train_csr(train)
        _l.check(self.lib.fb_diff_set_spike_train(self.h, int(ptr.size - 1), _l.ptr(ptr),
                                                   _l.ptr(idx) if idx.size else None))

    def forward(self, steps: int, seed=1, record: bool = True, spikeTrain=None) -> float:
        """forward(steps, {seed, record, spikeTrain}) (src/lifdiff.js:219): the
        train applies to this pass only, as in the reference."""
        self.set_spike_train(spikeTrain)
        L = ctypes.c_double()
        _l.check(self.lib.fb_diff_forward(self.h, int(steps), float(seed), 1 if record else 0, ctypes.byref(L)))
        self.steps = int(steps)
        return float(L.value)

    def backward(self, dLdSpike=None, truncate: int = 0, lossFrom: int | None = None, edges: bool = True) -> dict:
        if dLdSpike is not None or lossFrom is not None:
            self.set_loss(dLdSpike, lossFrom or 0)
        return self._grads(lambda *o: self.lib.fb_diff_backward(self.h, int(truncate), *o), edges, truncate)

    def grads(self, edges: bool = True) -> dict:
        """The results of the last backward pass (this network's or its batch's)."""
        return self._grads(lambda *o: self.lib.fb_diff_grads(self.h, *o), edges, 0)

    def _grads(self, call, edges, truncate):
        g = np.zeros(9, np.float64)
        lg = np.zeros(self.N, np.float32)
        ge = np.zeros(max(1, self.E), np.float32) if edges else None
        gi = np.zeros(self.N, np.float32)
        cl = np.zeros(3, np.int32)
        _l.check(call(_l.ptr(g), _l.ptr(lg), _l.ptr(ge) if ge is not None else None, _l.ptr(gi), _l.ptr(cl)))
        return {"params": dict(zip(DIFF_PARAMS, g.tolist())), "logGain": lg,
                "edge": ge[: self.E] if ge is not None else None, "inScale": gi, "clamped": cl,
                "truncate": int(truncate)}

    def type_grads(self, type_of, n_types: int) -> np.ndarray:
        t = np.ascontiguousarray(type_of, dtype=np.int32)
        out = np.zeros(int(n_types), np.float64)
        _l.check(self.lib.fb_diff_type_grads(self.h, _l.ptr(t), int(n_types), _l.ptr(out)))
        return out

    # ---------------------------------------------------------------- readback
    def state(self, name: str) -> np.ndarray:
        out = np.zeros(self.N, np.float32)
        _l.check(self.lib.fb_diff_get_state(self.h, STATE.index(name), _l.ptr(out)))
        return out

    def spikes(self):
        """(offsets[steps+1], idx, amp, forced) of the tape, each step's list
        in lif.c's order (forced spikes in drivenSet order, then ascending)."""
        off = np.zeros(self.steps + 1, np.int64)
        tot = ctypes.c_int64()
        _l.check(self.lib.fb_diff_spikes(self.h, _l.ptr(off), None, None, 0, ctypes.byref(tot)))
        n = int(tot.value)
        idx = np.zeros(max(1, n), np.int32)
        amp = np.zeros(max(1, n), np.float32)
        _l.check(self.lib.fb_diff_spikes(self.h, _l.ptr(off), _l.ptr(idx), _l.ptr(amp), n, ctypes.byref(tot)))
        u = idx[:n].view(np.uint32)
        ids = (u & 0x7FFFFFFF).astype(np.int32)
        forced = (u >> 31).astype(bool)
        amp = amp[:n]
        # canonical order inside each step: forced first (tape order), then by index
        order = np.lexsort((np.where(forced, 0, ids), ~forced, np.repeat(np.arange(self.steps), np.diff(off))))
        return off, ids[order], amp[order], forced[order]

    def readout(self):
        m = self.steps * self.R
        s = np.zeros(max(1, m), np.float32)
        r = np.zeros(max(1, m), np.float64)
        _l.check(self.lib.fb_diff_readout(self.h, _l.ptr(s), _l.ptr(r)))
        return s[:m].reshape(self.steps, self.R), r[:m].reshape(self.steps, self.R)

    @property
    def rng(self) -> int:
        out = np.zeros(1, np.uint32)
        _l.check(self.lib.fb_diff_rng(self.h, _l.ptr(out)))
        return int(out[0])

    def stat(self, name: str) -> float:
        v = ctypes.c_double()
        _l.check(self.lib.fb_diff_stat(self.h, name.encode(), ctypes.byref(v)))
        return float(v.value)


def autotune_checkpoint(N: int, steps: int, members: int = 1, budget_bytes: float = 200e9) -> int:
    """Checkpoint interval K for `members` concurrent instances of `steps`
    steps: minimises members * [(steps/K + 1) * S_ckpt + K * S_arena], the
    checkpoints (v, gE, gI, adapt, res, refr and the stream: 24 N + 4 bytes)
    against one segment's arena (v, gE, gI, adapt, spike amplitude and a mode
    byte: 21 N bytes per step), i.e. K = sqrt(steps * S_ckpt / S_arena).  The
    budget check counts two arenas, as a single network's backward keeps one
    segment's replay beside the sweep of the next.  Raises when the plan
    exceeds the budget."""
    import math
    s_ckpt = 24.0 * N + 4.0
    s_arena = 21.0 * N
    k = max(1, min(steps, int(round(math.sqrt(max(1, steps) * s_ckpt / s_arena)))))
    need = members * ((steps / k + 1) * s_ckpt + 2 * k * s_arena)
    if need > budget_bytes:
        raise ValueError(f"checkpoint plan needs {need / 1e9:.1f} GB > budget {budget_bytes / 1e9:.1f} GB")
    return k


class LIFDiffBatch:
    """B networks over one graph, each with its own inputs (gains, drive,
    seed, window), stepped in lockstep: every step's kernels run once for all
    members (the member is the launch's second grid dimension).
    ``type_grads`` reduces the members' type-gain gradients into one shared
    vector."""

    def __init__(self, N, indptr, indices, counts, sign, opts: dict | None, B: int):
        self.lib = _l.load()
        self.members = [LIFDiffHip(N, indptr, indices, counts, sign, opts) for _ in range(int(B))]
        hs = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _l.check(self.lib.fb_diff_batch_create(hs, len(self.members), ctypes.byref(h)))
        self.h = h

    def __len__(self):
        return len(self.members)

    def __getitem__(self, b: int) -> LIFDiffHip:
        return self.members[b]

    def forward(self, steps, seeds, record: bool = True) -> np.ndarray:
        st = np.ascontiguousarray(steps, dtype=np.int32)
        sd = np.ascontiguousarray(seeds, dtype=np.float64)
        L = np.zeros(len(self.members), np.float64)
        _l.check(self.lib.fb_diff_batch_forward(self.h, _l.ptr(st), _l.ptr(sd), 1 if record else 0, _l.ptr(L)))
        for m, n in zip(self.members, st.tolist()):
            m.steps = int(n)
        return L

    def backward(self, truncate: int = 0, edges: bool = False) -> list:
        _l.check(self.lib.fb_diff_batch_backward(self.h, int(truncate)))
        return [m.grads(edges=edges) for m in self.members]

    def run(self, steps, seeds, edges: bool = False, truncate: int = 0):
        """forward then backward for every member; [(loss, grads)] in member order."""
        L = self.forward(steps, seeds)
        return list(zip(L.tolist(), self.backward(truncate=truncate, edges=edges)))

    def type_grads(self, type_of, n_types: int) -> np.ndarray:
        """Sum over members of dL_b / d gain_T (each member's own gains)."""
        out = np.zeros(int(n_types), np.float64)
        for m in self.members:
            out += m.type_grads(type_of, n_types)
        return out

    def stat(self, name: str) -> float:
        v = ctypes.c_double()
        _l.check(self.lib.fb_diff_batch_stat(self.h, name.encode(), ctypes.byref(v)))
        return float(v.value)

    def destroy(self):
        if getattr(self, "h", None):
            _l.check(self.lib.fb_diff_batch_destroy(self.h))
            self.h = None
        for m in self.members:
            m.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
