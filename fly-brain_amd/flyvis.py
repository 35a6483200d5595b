"""The flyvis optic-lobe rate network on the GPU, with the host class of the
reference (src/flyvis.js:3-30): one instance per eye, luminance in on the
photoreceptor columns, graded activity out.

``parse_flyvis`` reads public/vision/flyvis.bin with its JSON sidecars
(src/flyvis.js:26-30); ``FlyVisHip`` is ``FlyVis`` over fb_fv_* (fv_step of
src/wasm/lif.c:82-93 as a gather kernel, bit-identical to the reference)."""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import lib as _l


def parse_flyvis(vision_dir: str) -> dict:
    """flyvis.bin (layout in flyvis.json: tid u8[N], u i16[N], v i16[N],
    bias f32[N], tau f32[N], indptr i32[N+1] by source, target i32[E],
    weight f32[E]) with the input map of flyvis_inputs.json; dt 0.02."""
    with open(os.path.join(vision_dir, "flyvis.json")) as f:
        meta = json.load(f)
    with open(os.path.join(vision_dir, "flyvis_inputs.json")) as f:
        inputs = json.load(f)
    N, E = int(meta["N"]), int(meta["E"])
    raw = np.fromfile(os.path.join(vision_dir, "flyvis.bin"), dtype=np.uint8)
    o = 0

    def take(dt, n):
        nonlocal o
        t = np.dtype(dt)
        a = raw[o:o + n * t.itemsize].view(t).copy()
        o += n * t.itemsize
        return a
    tid, u, v = take(np.uint8, N), take(np.int16, N), take(np.int16, N)
    bias, tau = take(np.float32, N), take(np.float32, N)
    indptr, target, weight = take(np.int32, N + 1), take(np.int32, E), take(np.float32, E)
    return {"N": N, "E": E, "tid": tid, "u": u, "v": v, "bias": bias, "tau": tau, "indptr": indptr,
            "target": target, "weight": weight, "types": meta["types"],
            "inputIdx": {k: np.asarray(x, dtype=np.int64) for k, x in inputs.items()}, "dt": 0.02}


class FlyVisModel:
    """The device copy of a model, shared by its instances (FlyVis's
    ``sharedParts``)."""

    def __init__(self, model: dict):
        self.lib = _l.load()
        self.N, self.E = int(model["N"]), int(model["E"])
        ip = np.ascontiguousarray(model["indptr"], dtype=np.int32)
        tg = np.ascontiguousarray(model["target"], dtype=np.int32)
        w = np.ascontiguousarray(model["weight"], dtype=np.float32)
        b = np.ascontiguousarray(model["bias"], dtype=np.float32)
        ta = np.ascontiguousarray(model["tau"], dtype=np.float32)
        h = ctypes.c_void_p()
        _l.check(self.lib.fb_fv_model_create(self.N, self.E, _l.ptr(ip), _l.ptr(tg), _l.ptr(w), _l.ptr(b),
                                             _l.ptr(ta), float(model.get("dt") or 0.02), ctypes.byref(h)))
        self.h = h
        self.bias = b
        self.inputIdx = model.get("inputIdx") or {}
        self.nCol = 0                      # set with the input map

    def release(self):
        if self.h:
            _l.check(self.lib.fb_fv_model_release(self.h))
            self.h = None


class FlyVisHip:
    """``new FlyVis(instance, memory, base, model)`` (src/flyvis.js:4-19) on
    the GPU: ``setInput(lum)``, ``step()``, ``reset()``, and the activity
    ``v`` (and the step's synaptic input ``acc``) as host copies.  Pass
    ``shared`` (another instance's ``model``) for the second eye."""

    def __init__(self, model: dict | None = None, shared: FlyVisModel | None = None):
        if shared is None:
            shared = FlyVisModel(model)
            self._own = True
        else:
            self._own = False
        self.model = shared
        self.lib = shared.lib
        self.N = shared.N
        self.inputIdx = shared.inputIdx
        h = ctypes.c_void_p()
        _l.check(self.lib.fb_fv_create(shared.h, ctypes.byref(h)))
        self.h = h
        self.x = np.zeros(self.N, dtype=np.float32)

    def setInput(self, lum):
        """Luminance per hex column, the same for R1-R8 (src/flyvis.js:21)."""
        lum = np.asarray(lum, dtype=np.float32)
        for ix in self.inputIdx.values():
            self.x[ix] = lum[: ix.size]
        _l.check(self.lib.fb_fv_set_input(self.h, _l.ptr(self.x)))

    def set_x(self, x):
        """The input of every node directly."""
        self.x = np.ascontiguousarray(x, dtype=np.float32).copy()
        _l.check(self.lib.fb_fv_set_input(self.h, _l.ptr(self.x)))

    def step(self, n: int = 1):
        _l.check(self.lib.fb_fv_step(self.h, int(n)))

    def reset(self):
        _l.check(self.lib.fb_fv_reset(self.h))

    def _get(self, which: int) -> np.ndarray:
        out = np.empty(self.N, dtype=np.float32)
        _l.check(self.lib.fb_fv_get(self.h, which, _l.ptr(out)))
        return out

    @property
    def v(self) -> np.ndarray:
        return self._get(0)

    @property
    def acc(self) -> np.ndarray:
        return self._get(1)

    def set_v(self, v):
        a = np.ascontiguousarray(v, dtype=np.float32)
        _l.check(self.lib.fb_fv_set_v(self.h, _l.ptr(a)))

    def destroy(self):
        if self.h:
            _l.check(self.lib.fb_fv_destroy(self.h))
            self.h = None
        if self._own and self.model is not None:
            self.model.release()
            self.model = None


class FlyVisDiffHip:
    """``new FlyVisDiff(model, shared)`` (src/flyvisdiff.js:36-185) on the GPU:
    the flyvis step as FlyVisHip runs it, a tape of every taped step in HBM
    and the reverse sweep.  ``shared``: the first eye's ``sharedParts`` (its
    device model), which is how the second eye is built.  ``model["kdt"]``,
    when present, replaces kdt = dt / max(tau, dt)."""

    def __init__(self, model: dict | None = None, shared: FlyVisModel | None = None):
        if shared is None:
            shared = FlyVisModel(model)
            lib = shared.lib
            if model.get("kdt") is not None:
                k = np.ascontiguousarray(model["kdt"], dtype=np.float32)
                _l.check(lib.fb_fv_model_set_kdt(shared.h, _l.ptr(k)))
            ix, col = [], []
            for a in shared.inputIdx.values():        # insertion order, as a JS object's keys
                ix += np.asarray(a).tolist()
                col += list(range(len(a)))
            if ix:
                nd = np.asarray(ix, dtype=np.int32)
                cl = np.asarray(col, dtype=np.int32)
                _l.check(lib.fb_fv_model_set_input_map(shared.h, int(nd.size), _l.ptr(nd), _l.ptr(cl)))
            shared.nCol = max(col) + 1 if col else 0
            self._own = True
        else:
            self._own = False
        self.model = shared
        self.lib = shared.lib
        self.N, self.E, self.nCol = shared.N, shared.E, shared.nCol
        self.dt = 0.02 if model is None else (model.get("dt") or 0.02)
        h = ctypes.c_void_p()
        _l.check(self.lib.fb_fv_create(shared.h, ctypes.byref(h)))
        self.h = h
        self.taping = False

    @property
    def sharedParts(self) -> FlyVisModel:
        return self.model

    def reset(self):
        _l.check(self.lib.fb_fv_reset(self.h))

    def setInput(self, lum):
        """lum[nCol]: luminance per hex column, the same for R1-R8."""
        a = np.ascontiguousarray(lum, dtype=np.float32)
        if a.size < self.nCol:
            raise ValueError(f"setInput: {a.size} columns, the map reads {self.nCol}")
        _l.check(self.lib.fb_fv_set_lum(self.h, _l.ptr(a)))

    def step(self, n: int = 1):
        _l.check(self.lib.fb_fv_step(self.h, int(n)))

    def settle(self, lum, steps: int = 150):
        """reset, hold ``lum`` and run ``steps`` untaped steps (src/flyvisdiff.js:98)."""
        self.reset()
        self.setInput(lum)
        self.step(steps)

    def startTape(self, steps: int):
        _l.check(self.lib.fb_fv_tape_start(self.h, int(steps)))
        self.taping = True

    def stepTaped(self, n: int = 1):
        _l.check(self.lib.fb_fv_step_taped(self.h, int(n)))

    @property
    def tape_len(self) -> int:
        n = ctypes.c_int32()
        _l.check(self.lib.fb_fv_tape_len(self.h, ctypes.byref(n)))
        return int(n.value)

    def backward(self, dLdV, g: dict | None = None, dLdLum=None) -> dict:
        """dLdV [steps*N]; ``g`` = {weight[E], bias[N], kdt[N]} float32
        accumulators, added to in place (created when absent); ``dLdLum``
        [steps*nCol] float32 accumulator, added to in place, or None."""
        n = self.tape_len
        d = np.ascontiguousarray(dLdV, dtype=np.float32).reshape(-1)
        if d.size != n * self.N:
            raise ValueError(f"backward: dLdV holds {d.size} values, the tape {n} steps of {self.N}")
        g = g or {"weight": np.zeros(self.E, np.float32), "bias": np.zeros(self.N, np.float32),
                  "kdt": np.zeros(self.N, np.float32)}
        for k, size in (("weight", self.E), ("bias", self.N), ("kdt", self.N)):
            a = g[k]
            if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous and a.size == size):
                raise ValueError(f"backward: g[{k!r}] must be a contiguous float32 array of {size}")
        lp = None
        if dLdLum is not None:
            if not (isinstance(dLdLum, np.ndarray) and dLdLum.dtype == np.float32 and dLdLum.flags.c_contiguous
                    and dLdLum.size == n * self.nCol):
                raise ValueError(f"backward: dLdLum must be a contiguous float32 array of {n * self.nCol}")
            lp = _l.ptr(dLdLum)
        _l.check(self.lib.fb_fv_backward(self.h, _l.ptr(d), _l.ptr(g["weight"]), _l.ptr(g["bias"]),
                                         _l.ptr(g["kdt"]), lp))
        return g

    def _get(self, which: int) -> np.ndarray:
        out = np.empty(self.N, dtype=np.float32)
        _l.check(self.lib.fb_fv_get(self.h, which, _l.ptr(out)))
        return out

    @property
    def v(self) -> np.ndarray:
        return self._get(0)

    def set_v(self, v):
        a = np.ascontiguousarray(v, dtype=np.float32)
        _l.check(self.lib.fb_fv_set_v(self.h, _l.ptr(a)))

    def destroy(self):
        if self.h:
            _l.check(self.lib.fb_fv_destroy(self.h))
            self.h = None
        if self._own and self.model is not None:
            self.model.release()
            self.model = None
