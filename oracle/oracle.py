"""ctypes front end for the CPU checkers.  TEST INFRASTRUCTURE ONLY.

``OracleBrain(kind="oracle")`` drives the restatement in lif_oracle.c;
``kind="ref"`` drives the reference's own lif.c compiled into oracle/_ref/
(present only where /root/reference was available at build time).  Both
expose the same calls, so a test can run one against the other.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblif_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_lif.so")

FIELDS = ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay",
          "traceTau", "adaptInc", "adaptTau", "depU", "depTau", "coba", "eExc",
          "eInh", "bgRate", "bgAmp", "preInh")
ARRAYS = ("v", "gE", "gI", "refr", "trace", "adapt", "res", "bias", "thr", "drive")


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in FIELDS]


def params_struct(p: dict) -> _Params:
    s = _Params()
    for n in FIELDS:
        v = p.get(n, 0.0) or 0.0
        setattr(s, n, float(bool(v)) if n == "coba" else float(v))
    return s


_libs: dict = {}


def available(kind: str) -> bool:
    return os.path.exists(ORACLE_SO if kind == "oracle" else REF_SO)


def _lib(kind: str):
    if kind in _libs:
        return _libs[kind]
    path = ORACLE_SO if kind == "oracle" else REF_SO
    lib = ctypes.CDLL(path)
    pre = "fbo_" if kind == "oracle" else "ref_"
    P, I32, I64, U32, F = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_float
    sig = {
        "create": (P, [I32, I64, P, P, P, P, ctypes.POINTER(_Params), ctypes.c_double]),
        "destroy": (None, [P]), "reset": (None, [P]),
        "set_driven": (None, [P, P, I32]),
        "array": (ctypes.POINTER(F), [P, I32]),
        "spike_count": (ctypes.POINTER(U32), [P]),
        "rng_state": (U32, [P]), "set_rng": (None, [P, U32]), "head": (I32, [P]), "nslots_of": (I32, [P]),
        "bg_acc": (F, [P]),
        "last_fired": (ctypes.POINTER(I32), [P, ctypes.POINTER(I32)]),
        "step": (I32, [P]), "run": (I64, [P, I32, P, P, I64]),
        "add": (None, [P, I32, I32, ctypes.c_double]),
        "store": (None, [P, I32, I32, ctypes.c_double]),
        "set_background": (None, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, pre + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    if kind == "oracle":
        lib.fbo_seed_hash.restype = U32
        lib.fbo_seed_hash.argtypes = [ctypes.c_double]
        lib.fbo_xorshift32.restype = U32
        lib.fbo_xorshift32.argtypes = [ctypes.POINTER(U32)]
        lib.fbo_nslots.restype = I32
        lib.fbo_nslots.argtypes = [ctypes.c_double, ctypes.c_double]
        fns["_lib"] = lib
    _libs[kind] = fns
    return fns


def seed_hash(seed) -> int:
    return int(_lib("oracle")["_lib"].fbo_seed_hash(float(seed)))


def xorshift_stream(state: int, n: int) -> np.ndarray:
    st = ctypes.c_uint32(state)
    f = _lib("oracle")["_lib"].fbo_xorshift32
    return np.array([f(ctypes.byref(st)) for _ in range(n)], dtype=np.uint32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleBrain:
    """One brain on the CPU checker: same state arrays and setters as the
    reference's LIFWasm (src/lifwasm.js:36-87)."""

    def __init__(self, graph, params: dict, seed=1, kind: str = "oracle", W=None, sign=None):
        self.kind = kind
        self.f = _lib(kind)
        self.p = dict(params)
        self.N = int(graph.N)
        self._indptr = np.ascontiguousarray(graph.indptr, dtype=np.uint32)
        self._indices = np.ascontiguousarray(graph.indices, dtype=np.uint32)
        self._W = np.ascontiguousarray(graph.weights if W is None else W, dtype=np.float32)
        self._sign = np.ascontiguousarray(graph.sign if sign is None else sign, dtype=np.float32)
        ps = params_struct(self.p)
        self.h = self.f["create"](self.N, int(self._indices.size), _ptr(self._indptr), _ptr(self._indices),
                                  _ptr(self._W), _ptr(self._sign), ctypes.byref(ps), float(seed))
        self.drivenSet: list = []
        self._driven_dirty = True

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.f["destroy"](h)
            self.h = None

    # ---------------------------------------------------------------- state
    def array(self, name: str) -> np.ndarray:
        ptr = self.f["array"](self.h, ARRAYS.index(name))
        return np.ctypeslib.as_array(ptr, shape=(self.N,))

    def __getattr__(self, name):
        if name in ARRAYS:
            return self.array(name)
        raise AttributeError(name)

    @property
    def spikeCount(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.f["spike_count"](self.h), shape=(self.N,))

    @property
    def rng(self) -> int:
        return int(self.f["rng_state"](self.h))

    @property
    def head(self) -> int:
        return int(self.f["head"](self.h))

    @property
    def nslots(self) -> int:
        return int(self.f["nslots_of"](self.h))

    def state_digest(self) -> dict:
        return {n: self.array(n).copy() for n in ARRAYS[:7]} | {"spikeCount": self.spikeCount.copy()}

    # ---------------------------------------------------------------- setters (lifwasm.js:70-78)
    def setDriveOne(self, i: int, rate: float):
        self.f["store"](self.h, ARRAYS.index("drive"), int(i), float(rate))
        if rate > 0:
            if i not in self.drivenSet:
                self.drivenSet.append(int(i))
                self._driven_dirty = True
        elif i in self.drivenSet:
            self.drivenSet.remove(int(i))
            self._driven_dirty = True

    def setDrive(self, ix, rate: float):
        for i in ix:
            self.setDriveOne(int(i), rate)

    def setBias(self, ix, mv: float):
        for i in ix:
            self.f["store"](self.h, ARRAYS.index("bias"), int(i), float(mv))

    def setThr(self, i: int, mv: float):
        self.f["store"](self.h, ARRAYS.index("thr"), int(i), float(mv))

    def addG(self, i: int, e: float, ii: float = 0.0):
        self.f["add"](self.h, ARRAYS.index("gE"), int(i), float(e))
        self.f["add"](self.h, ARRAYS.index("gI"), int(i), float(ii or 0.0))

    def pulse(self, ix, mv: float):
        for i in ix:
            self.f["add"](self.h, ARRAYS.index("gE"), int(i), float(mv))

    def setBackground(self, rate_hz: float, amp_mv: float):
        self.f["set_background"](self.h, float(rate_hz), float(amp_mv), float(self.p["dt"]))

    def reset(self):
        self.f["reset"](self.h)

    # ---------------------------------------------------------------- stepping
    def _sync_driven(self):
        if self._driven_dirty:
            lst = np.array(self.drivenSet, dtype=np.int32)
            self.f["set_driven"](self.h, _ptr(lst), int(lst.size))
            self._driven_dirty = False

    def step(self) -> np.ndarray:
        self._sync_driven()
        self.f["step"](self.h)
        n = ctypes.c_int32()
        ptr = self.f["last_fired"](self.h, ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()

    def run(self, steps: int, raster_cap: int | None = None):
        """Advance ``steps`` steps; returns (nfired per step, concatenated spike lists)."""
        self._sync_driven()
        nf = np.zeros(steps, dtype=np.int32)
        cap = raster_cap if raster_cap is not None else steps * 64
        ras = np.zeros(max(1, cap), dtype=np.int32)
        tot = self.f["run"](self.h, steps, _ptr(nf), _ptr(ras), cap)
        if tot > cap:
            ras = None
        else:
            ras = ras[:tot].copy()
        return nf, ras, int(tot)


# ---------------------------------------------------------------- differentiable model

DIFF_FIELDS = ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay",
               "wSyn", "adaptInc", "adaptTau", "depU", "depTau",
               "eExc", "eInh", "inhGain", "kcThreshold", "laminaBias",
               "surrogateBeta", "adjClip", "sizeAlpha", "maxSizeScale", "boostCap")
DIFF_INPUTS = ("outScale", "logGain", "thrOffset", "sizeLog", "drive", "inScale", "surrogateBeta")
DIFF_STATE = ("v", "gE", "gI", "adapt", "res", "refr", "spikeCount", "inScale")
DIFF_PARAMS = ("wSyn", "sizeAlpha", "inhGain", "eInh", "vThresh", "kcThreshold", "laminaBias",
               "adaptInc", "depU")


class _DiffParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in DIFF_FIELDS] + [("soft", ctypes.c_int32), ("checkpoint", ctypes.c_int32),
                                                              ("exact32", ctypes.c_int32)]


def _diff_lib():
    if "diff" in _libs:
        return _libs["diff"]
    lib = ctypes.CDLL(ORACLE_SO)
    P, I32, I64, F, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
    sig = {
        "fdo_create": (P, [I32, I64, P, P, P, P, ctypes.POINTER(_DiffParams)]),
        "fdo_destroy": (None, [P]),
        "fdo_set_input": (None, [P, I32, P]),
        "fdo_set_edge_gain": (None, [P, P]),
        "fdo_set_masks": (None, [P, P, P]),
        "fdo_set_driven": (None, [P, P, I32]),
        "fdo_set_loss": (None, [P, P, I32]),
        "fdo_set_readout": (None, [P, P, I32, P, D]),
        "fdo_forward": (D, [P, I32, D, I32]),
        "fdo_backward": (None, [P, I32, P, P, P, P, P]),
        "fdo_state": (ctypes.POINTER(F), [P, I32]),
        "fdo_rng": (ctypes.c_uint32, [P]),
        "fdo_spikes": (I64, [P, P, P, P, I64]),
        "fdo_readout": (I64, [P, P, P]),
        "fdo_set_drive_schedule": (None, [P, P, I32, I32, P, I32, I32]),
        "fdo_set_loss_per_epoch": (None, [P, P, I32, I32]),
        "fdo_drive_grad": (None, [P, P]),
        "fdo_set_spike_train": (None, [P, I32, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs["diff"] = lib
    return lib


class OracleDiff:
    """The differentiable model on the CPU checker: LIFDiff's forward with a
    tape, and its backward (src/lifdiff.js:84-686), plus the per-edge gradient
    and the filtered-rate readout loss.

    ``counts`` are the delivered synapse counts (gate folded in as zeros),
    ``sign`` is +1 / -1 / 0 per neuron; ``params`` holds LIFDiff's options."""

    def __init__(self, N, indptr, indices, counts, sign, params: dict):
        self._dsched = None
        self.lib = _diff_lib()
        self.N = int(N)
        self.p = dict(params)
        self._indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
        self._indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.E = int(self._indices.size)
        self._counts = np.ascontiguousarray(counts, dtype=np.float32)
        self._sign = np.ascontiguousarray(sign, dtype=np.float32)
        ps = _DiffParams()
        for n in DIFF_FIELDS:
            setattr(ps, n, float(self.p[n]))
        ps.soft = 1 if self.p["soft"] else 0
        ps.checkpoint = int(self.p["checkpoint"])
        ps.exact32 = 1 if self.p.get("exact32") else 0
        self.h = self.lib.fdo_create(self.N, self.E, _ptr(self._indptr), _ptr(self._indices), _ptr(self._counts),
                                     _ptr(self._sign), ctypes.byref(ps))
        self._keep = []

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.fdo_destroy(h)
            self.h = None

    def set_input(self, name: str, values):
        """A per-neuron input; None clears a caller-given inScale or the
        per-neuron surrogateBeta (back to the derived inScale, the scalar)."""
        if values is None:
            assert name in ("inScale", "surrogateBeta")
            self.lib.fdo_set_input(self.h, DIFF_INPUTS.index(name), None)
            return
        a = np.ascontiguousarray(values, dtype=np.float32)
        assert a.size == self.N
        self.lib.fdo_set_input(self.h, DIFF_INPUTS.index(name), _ptr(a))

    def set_edge_gain(self, eg):
        if eg is None:
            self.lib.fdo_set_edge_gain(self.h, None)
            return
        a = np.ascontiguousarray(eg, dtype=np.float32)
        assert a.size == self.E
        self.lib.fdo_set_edge_gain(self.h, _ptr(a))

    def set_masks(self, thr_mask=None, bias_mask=None):
        t = None if thr_mask is None else np.ascontiguousarray(thr_mask, dtype=np.uint8)
        b = None if bias_mask is None else np.ascontiguousarray(bias_mask, dtype=np.uint8)
        self.lib.fdo_set_masks(self.h, None if t is None else _ptr(t), None if b is None else _ptr(b))

    def set_driven(self, ix):
        a = np.ascontiguousarray(ix, dtype=np.int32)
        self.lib.fdo_set_driven(self.h, _ptr(a), int(a.size))

    def set_loss(self, dLdSpike, loss_from: int = 0):
        a = np.ascontiguousarray(dLdSpike, dtype=np.float32)
        self.lib.fdo_set_loss(self.h, _ptr(a), int(loss_from))

    def set_readout(self, idx, target_hz, tau_ms: float):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        t = np.ascontiguousarray(target_hz, dtype=np.float32)
        self.R = int(i.size)
        self.lib.fdo_set_readout(self.h, _ptr(i), self.R, _ptr(t), float(tau_ms))

    def set_drive_schedule(self, idx, epoch_steps: int, sched, drive_soft: bool = False):
        """driveIdx / driveEpoch with the rates sched[e][k] the onEpoch hook sets
        at the start of each epoch (src/lifdiff.js:135-144, 228-234); None clears."""
        if idx is None:
            self.lib.fdo_set_drive_schedule(self.h, None, 0, 1, None, 0, 1 if drive_soft else 0)
            self._dsched = None
            return
        i = np.ascontiguousarray(idx, dtype=np.int32)
        sc = np.ascontiguousarray(sched, dtype=np.float32).reshape(-1, i.size)
        self._dsched = (i, sc)
        self.lib.fdo_set_drive_schedule(self.h, _ptr(i), int(i.size), int(epoch_steps), _ptr(sc), int(sc.shape[0]),
                                        1 if drive_soft else 0)

    def set_loss_per_epoch(self, dper, epoch_steps: int):
        """dLdSpikePerEpoch (src/lifdiff.js:519-522): dper[e][i]; None clears."""
        if dper is None:
            self.lib.fdo_set_loss_per_epoch(self.h, None, 0, int(epoch_steps))
            return
        a = np.ascontiguousarray(dper, dtype=np.float32).reshape(-1, self.N)
        self.lib.fdo_set_loss_per_epoch(self.h, _ptr(a), int(a.shape[0]), int(epoch_steps))

    def drive_grad(self) -> np.ndarray:
        i, sc = self._dsched
        out = np.zeros(sc.shape, dtype=np.float32)
        self.lib.fdo_drive_grad(self.h, _ptr(out))
        return out

    def set_spike_train(self, train):
        """train[t]: the neurons forced at step t (None or empty for none); None clears."""
        if train is None:
            self.lib.fdo_set_spike_train(self.h, 0, None, None)
            return
        lists = [np.asarray(x if x is not None else [], dtype=np.int32).ravel() for x in train]
        ptr = np.zeros(len(lists) + 1, dtype=np.int64)
        ptr[1:] = np.cumsum([x.size for x in lists]) if lists else []
        idx = np.concatenate(lists).astype(np.int32) if lists and ptr[-1] else np.zeros(1, np.int32)
        self._train = (ptr, idx)
        self.lib.fdo_set_spike_train(self.h, len(lists), _ptr(ptr), _ptr(idx))

    def forward(self, steps: int, seed=1, record: bool = True, spikeTrain=None) -> float:
        """forward(steps, {seed, record, spikeTrain}) (src/lifdiff.js:219); the
        train applies to this pass, as in the reference."""
        self.set_spike_train(spikeTrain)
        self.steps = int(steps)
        return float(self.lib.fdo_forward(self.h, self.steps, float(seed), 1 if record else 0))

    def backward(self, truncate: int = 0) -> dict:
        g = np.zeros(9, dtype=np.float64)
        lg = np.zeros(self.N, dtype=np.float32)
        ge = np.zeros(max(1, self.E), dtype=np.float32)
        gi = np.zeros(self.N, dtype=np.float32)
        cl = np.zeros(3, dtype=np.int32)
        self.lib.fdo_backward(self.h, int(truncate), _ptr(g), _ptr(lg), _ptr(ge), _ptr(gi), _ptr(cl))
        return {"params": dict(zip(DIFF_PARAMS, g.tolist())), "logGain": lg, "edge": ge[: self.E],
                "inScale": gi, "clamped": cl}

    def state(self, name: str) -> np.ndarray:
        ptr = self.lib.fdo_state(self.h, DIFF_STATE.index(name))
        return np.ctypeslib.as_array(ptr, shape=(self.N,)).copy()

    @property
    def rng(self) -> int:
        return int(self.lib.fdo_rng(self.h))

    def spikes(self):
        """(offsets[steps+1], idx, amp, forced) of the recorded spike lists."""
        off = np.zeros(self.steps + 1, dtype=np.int64)
        tot = int(self.lib.fdo_spikes(self.h, _ptr(off), None, None, 0))
        idx = np.zeros(max(1, tot), dtype=np.int32)
        amp = np.zeros(max(1, tot), dtype=np.float32)
        self.lib.fdo_spikes(self.h, _ptr(off), _ptr(idx), _ptr(amp), tot)
        u = idx[:tot].view(np.uint32)
        return off, (u & 0x7FFFFFFF).astype(np.int32), amp[:tot], (u >> 31).astype(bool)

    def readout(self):
        m = self.steps * self.R
        s = np.zeros(max(1, m), dtype=np.float32)
        r = np.zeros(max(1, m), dtype=np.float64)
        self.lib.fdo_readout(self.h, _ptr(s), _ptr(r))
        return s[:m].reshape(self.steps, self.R), r[:m].reshape(self.steps, self.R)


# ------------------------------------------------------------------ LIFNetwork (the JS twin)

NET_FIELDS = ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay", "wSyn", "noise",
              "traceTau", "adaptInc", "adaptTau", "depU", "depTau", "minSyn", "preInh", "coba", "eExc",
              "eInh", "inhGain")
NET_ARRAYS = ("v", "gE", "gI", "refr", "trace", "adapt", "res", "bias", "thr", "drive", "weights")
# src/lif.js:6-15
NET_DEFAULTS = {"dt": 0.5, "vRest": -52.0, "vThresh": -45.0, "vReset": -52.0, "tauM": 20.0, "tauSyn": 5.0,
                "tRef": 2.2, "delay": 1.8, "wSyn": 0.275, "noise": 0.0, "traceTau": 30.0, "adaptInc": 2.0,
                "adaptTau": 100.0, "depU": 0.2, "depTau": 200.0, "minSyn": 1.0, "preInh": 0.0, "coba": False,
                "eExc": 0.0, "eInh": -70.0, "inhGain": 1.0}


class _NetParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in NET_FIELDS]


class OracleNet:
    """LIFNetwork (src/lif.js:16-105) restated in C with Math.random = mulberry32(rseed)."""

    def __init__(self, N, indptr, indices, weights, nt, params=None, inScale=None, outScale=None,
                 sensoryMask=None, ntSign=None, preSign=None, rseed: int = 1):
        lib = _lib_net()
        self.lib = lib
        self.N = int(N)
        p = dict(NET_DEFAULTS)
        p.update(params or {})
        self.p = p
        s = _NetParams()
        for n in NET_FIELDS:
            setattr(s, n, float(bool(p[n])) if n == "coba" else float(p[n]))
        self._keep = []

        def arr(a, dt):
            if a is None:
                return None
            x = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(x)
            return x.ctypes.data_as(ctypes.c_void_p)

        ip = np.ascontiguousarray(indptr, dtype=np.uint32)
        self.E = int(ip[-1])
        self.h = lib.lno_create(self.N, arr(ip, np.uint32), arr(indices, np.uint32), arr(weights, np.float32),
                                arr(nt, np.uint8), ctypes.byref(s), arr(inScale, np.float32),
                                arr(outScale, np.float32), arr(sensoryMask, np.uint8), arr(ntSign, np.float64),
                                arr(preSign, np.float64), ctypes.c_uint32(rseed & 0xFFFFFFFF))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.lno_destroy(self.h)
            self.h = None

    def _ix(self, ix):
        a = np.ascontiguousarray(ix, dtype=np.int32)
        return a, a.ctypes.data_as(ctypes.c_void_p)

    def reset(self):
        self.lib.lno_reset(self.h)

    def setBias(self, ix, mv):
        a, p = self._ix(ix)
        self.lib.lno_set_bias(self.h, p, int(a.size), float(mv))

    def setDrive(self, ix, rate):
        a, p = self._ix(ix)
        self.lib.lno_set_drive(self.h, p, int(a.size), float(rate))

    def setDriveOne(self, i, rate):
        self.lib.lno_set_drive_one(self.h, int(i), float(rate))

    def setThr(self, i, mv):
        self.lib.lno_set_thr(self.h, int(i), float(mv))

    def addG(self, i, e, ii=0.0):
        self.lib.lno_add_g(self.h, int(i), float(e), float(ii or 0.0))

    def pulse(self, ix, mv):
        a, p = self._ix(ix)
        self.lib.lno_pulse(self.h, p, int(a.size), float(mv))

    def step(self) -> np.ndarray:
        nf = int(self.lib.lno_step(self.h))
        return np.ctypeslib.as_array(self.lib.lno_fired(self.h), shape=(max(1, self.N),))[:nf].copy()

    def run(self, steps: int) -> np.ndarray:
        nf = np.zeros(max(1, steps), dtype=np.int32)
        self.lib.lno_run(self.h, int(steps), nf.ctypes.data_as(ctypes.c_void_p))
        return nf[:steps]

    def array(self, name: str) -> np.ndarray:
        n = self.E if name == "weights" else self.N
        ptr = self.lib.lno_array(self.h, NET_ARRAYS.index(name))
        return np.ctypeslib.as_array(ptr, shape=(max(1, n),))[:n].copy()

    @property
    def spikeCount(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.lib.lno_spike_count(self.h), shape=(max(1, self.N),))[: self.N].copy()

    @property
    def nAwake(self) -> int:
        return int(self.lib.lno_n_awake(self.h))

    @property
    def t(self) -> float:
        return float(self.lib.lno_time(self.h))

    @property
    def nslots(self) -> int:
        return int(self.lib.lno_nslots(self.h))


def _lib_net():
    lib = ctypes.CDLL(ORACLE_SO)
    P, I32, I64, D, U32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32
    sig = {
        "lno_create": (P, [I32, P, P, P, P, ctypes.POINTER(_NetParams), P, P, P, P, P, U32]),
        "lno_destroy": (None, [P]), "lno_reset": (None, [P]),
        "lno_set_bias": (None, [P, P, I32, D]), "lno_set_drive": (None, [P, P, I32, D]),
        "lno_set_drive_one": (None, [P, I32, D]), "lno_set_thr": (None, [P, I32, D]),
        "lno_add_g": (None, [P, I32, D, D]), "lno_pulse": (None, [P, P, I32, D]),
        "lno_step": (I32, [P]), "lno_fired": (ctypes.POINTER(I32), [P]),
        "lno_array": (ctypes.POINTER(ctypes.c_float), [P, I32]),
        "lno_spike_count": (ctypes.POINTER(U32), [P]), "lno_n_awake": (I32, [P]),
        "lno_time": (D, [P]), "lno_nslots": (I32, [P]), "lno_run": (I64, [P, I32, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


# ---------------------------------------------------------------- flyvis

def fv_step(kind: str, bias, kdt, indptr, target, weight, v, acc, x):
    """One fv_step on the checker: ``kind="oracle"`` runs the restatement
    (fbo_fv_step), ``kind="ref"`` the reference's own fv_step compiled into
    oracle/_ref (src/wasm/lif.c:86-93).  v and acc are updated in place
    (float32 arrays)."""
    lib = ctypes.CDLL(ORACLE_SO if kind == "oracle" else REF_SO)
    f = lib.fbo_fv_step if kind == "oracle" else lib.fv_step
    f.restype = None
    P = ctypes.c_void_p
    f.argtypes = [ctypes.c_int32, P, P, P, P, P, P, P, P]
    arrs = [np.ascontiguousarray(a, dtype=dt) for a, dt in
            ((bias, np.float32), (kdt, np.float32), (indptr, np.int32), (target, np.int32), (weight, np.float32),
             (x, np.float32))]
    assert v.dtype == np.float32 and acc.dtype == np.float32 and v.flags.c_contiguous and acc.flags.c_contiguous
    f(int(v.size), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3]), _ptr(arrs[4]), _ptr(v), _ptr(acc),
      _ptr(arrs[5]))


def fv_kdt(tau, dt: float = 0.02) -> np.ndarray:
    """kdt = dt / max(tau, dt) in binary64, stored as binary32 (src/flyvis.js:13)."""
    t = np.asarray(tau, dtype=np.float32).astype(np.float64)
    return (dt / np.maximum(t, dt)).astype(np.float32)


class FlyVisDiffOracle:
    """FlyVisDiff (src/flyvisdiff.js:36-185) on the CPU checker: a taped forward
    (fv_step, with each step's pre-update activity and integrator argument
    kept) and the reverse sweep.  ``model`` as parse_flyvis returns it; the
    input map is flattened type by type as the reference does."""

    def __init__(self, model: dict, kdt=None):
        self.m = model
        self.N, self.E = int(model["N"]), int(model["E"])
        self.bias = np.ascontiguousarray(model["bias"], dtype=np.float32)
        self.kdt = np.ascontiguousarray(kdt if kdt is not None else fv_kdt(model["tau"], model.get("dt") or 0.02),
                                        dtype=np.float32)
        self.indptr = np.ascontiguousarray(model["indptr"], dtype=np.int32)
        self.target = np.ascontiguousarray(model["target"], dtype=np.int32)
        self.weight = np.ascontiguousarray(model["weight"], dtype=np.float32)
        ix, col = [], []
        for a in (model.get("inputIdx") or {}).values():
            ix += list(np.asarray(a).tolist())
            col += list(range(len(a)))
        self.inNode = np.asarray(ix, dtype=np.int32)
        self.inCol = np.asarray(col, dtype=np.int32)
        self.nCol = int(self.inCol.max()) + 1 if self.inCol.size else 0
        self.v = self.bias.copy()
        self.acc = np.zeros(self.N, np.float32)
        self.x = np.zeros(self.N, np.float32)
        lib = ctypes.CDLL(ORACLE_SO)
        self._step = lib.fbo_fv_step
        self._step.restype = None
        self._back = lib.fbo_fv_back_step
        self._back.restype = None
        self.tape = None

    def reset(self):
        self.v = self.bias.copy()

    def setInput(self, lum):
        lum = np.asarray(lum, dtype=np.float32)
        self.x[self.inNode] = lum[self.inCol]
        self.lum = lum.copy()

    def set_x(self, x):
        self.x = np.ascontiguousarray(x, dtype=np.float32).copy()

    def step(self):
        vPre = self.v.copy()
        self._step(ctypes.c_int32(self.N), _ptr(self.bias), _ptr(self.kdt), _ptr(self.indptr), _ptr(self.target),
                   _ptr(self.weight), _ptr(self.v), _ptr(self.acc), _ptr(self.x))
        if self.tape is not None:
            u = ((-vPre + self.bias) + self.acc) + self.x       # binary32, as the step formed it
            self.tape.append((vPre, u.astype(np.float32), getattr(self, "lum", None)))

    def start_tape(self):
        self.tape = []

    def backward(self, dLdV, g=None, dLdLum=None):
        """dLdV [steps][N]; g = {weight, bias, kdt} accumulators (f32); dLdLum
        [steps][nCol] accumulator or None."""
        N = self.N
        g = g or {"weight": np.zeros(self.E, np.float32), "bias": np.zeros(N, np.float32),
                  "kdt": np.zeros(N, np.float32)}
        lv = np.zeros(N, np.float32)
        lacc = np.zeros(N, np.float32)
        dLdV = np.ascontiguousarray(dLdV, dtype=np.float32).reshape(len(self.tape), N)
        for t in range(len(self.tape) - 1, -1, -1):
            vPre, u, _ = self.tape[t]
            row = None if dLdLum is None else dLdLum[t]
            self._back(ctypes.c_int32(N), _ptr(self.indptr), _ptr(self.target), _ptr(self.weight), _ptr(self.kdt),
                       _ptr(vPre), _ptr(u), _ptr(dLdV[t]), _ptr(lv), _ptr(lacc), _ptr(g["weight"]), _ptr(g["bias"]),
                       _ptr(g["kdt"]), None if row is None else _ptr(row), _ptr(self.inNode), _ptr(self.inCol),
                       ctypes.c_int32(self.inNode.size))
        return g


class VisualChainOracle:
    """VisualChain (src/visdiff.js:70-258) on the CPU checker: two
    FlyVisDiffOracle eyes, the coupling restated pair by pair in binary64 with
    the reference's binary32 stores, and OracleDiff as the CNS.  The rates
    every epoch start sets depend on the optic lobe alone, so they are
    computed eye step by eye step first and given to the CNS as its drive
    schedule -- the sequence of setDriveOne calls the onEpoch hook makes.

    ``cns``: {N, indptr, indices, counts, sign} (counts before the gate);
    ``params``: LIFDiff's options, complete; ``coupling``: gain, dead, cap and
    "coupling" ("hard" | "soft")."""

    GATE_K, CAP_KNEE = 8000.0, 0.05            # src/visdiff.js:68

    def __init__(self, cns, fv_model, fv_map, params, coupling=None, fvDt=None, driveSoft=False,
                 sensoryMask=None):
        import math
        self.c = {"gain": 250.0, "dead": 0.02, "cap": 200.0, "coupling": "hard", **(coupling or {})}
        self.eyes = [FlyVisDiffOracle(fv_model, kdt=fv_model.get("kdt")),
                     FlyVisDiffOracle(fv_model, kdt=fv_model.get("kdt"))]
        self.fvN, self.fvE, self.nCol = self.eyes[0].N, self.eyes[0].E, self.eyes[0].nCol
        self.dt_fv = float(fv_model.get("dt") or 0.02)
        self.pairs = [[(int(q[0]), int(q[1])) for q in fv_map["eyes"][sd]["pairs"]] for sd in ("L", "R")]
        idx = [n for P in self.pairs for (n, _) in P]
        if len(set(idx)) != len(idx):
            raise ValueError("a CNS neuron is driven by both eyes")
        self.driveIdx = np.asarray(idx, dtype=np.int32)
        self.nD = len(idx)
        self.slot = {n: k for k, n in enumerate(idx)}
        N = int(cns["N"])
        sensory = np.zeros(N, np.uint8) if sensoryMask is None else np.asarray(sensoryMask, np.uint8).copy()
        sensory[self.driveIdx] = 1              # src/visdiff.js:107-109
        counts = np.asarray(cns["counts"], np.float32).copy()
        ind = np.asarray(cns["indices"])
        for j in range(counts.size):            # LIFDiff's gate (src/lifdiff.js:167-172)
            if counts[j] < params["minSyn"] or sensory[ind[j]]:
                counts[j] = 0.0
        self.net = OracleDiff(N, cns["indptr"], cns["indices"], counts, cns["sign"], params)
        fv_ms = fvDt if fvDt is not None else self.dt_fv * 1000.0
        self.epoch = max(1, int(math.floor(fv_ms / params["dt"] + 0.5)))   # Math.round
        self.driveSoft = bool(driveSoft)
        self.vRest = None
        self.eyeV0 = None

    # the coupling (src/visdiff.js:131-159), one deviation at a time
    def _gate(self, a):
        """1 / (1 + exp(-K (a - dead))), which is 0 once exp overflows to
        Infinity (past ~709.78), as in JavaScript"""
        import math
        try:
            e = math.exp(-self.GATE_K * (a - self.c["dead"]))
        except OverflowError:
            return 0.0
        return 1.0 / (1.0 + e)

    def _rate(self, a):
        import math
        c = self.c
        if c["coupling"] != "soft":
            if a <= c["dead"]:
                return 0.0
            return min(c["cap"], c["gain"] * a)
        gate = self._gate(a)
        u = c["cap"] - c["gain"] * a
        inner = c["cap"] - 0.5 * (u + math.sqrt(u * u + self.CAP_KNEE ** 2))
        return max(0.0, inner * gate)

    def _drate(self, a):
        import math
        c = self.c
        if c["coupling"] != "soft":
            return 0.0 if (a <= c["dead"] or c["gain"] * a >= c["cap"]) else c["gain"]
        u = c["cap"] - c["gain"] * a
        inner = c["cap"] - 0.5 * (u + math.sqrt(u * u + self.CAP_KNEE ** 2))
        d_inner = 0.5 * (1 + u / math.sqrt(u * u + self.CAP_KNEE ** 2))
        gate = self._gate(a)
        return c["gain"] * d_inner * gate + inner * gate * (1 - gate) * self.GATE_K

    def _drate_gain(self, a):
        import math
        c = self.c
        if c["coupling"] != "soft":
            return 0.0 if (a <= c["dead"] or c["gain"] * a >= c["cap"]) else a
        u = c["cap"] - c["gain"] * a
        d_inner = 0.5 * (1 + u / math.sqrt(u * u + self.CAP_KNEE ** 2))
        gate = self._gate(a)
        return a * d_inner * gate

    def settle(self, level=0.5, steps=150):
        grey = np.full(self.nCol, level, np.float32)
        for e in self.eyes:
            e.tape = None
            e.reset()
            e.setInput(grey)
            for _ in range(steps):
                e.step()
        self.vRest = self.eyes[0].v.copy()
        self.eyeV0 = [e.v.copy() for e in self.eyes]
        return self.vRest

    def forward(self, steps, seed=1, lum=None, record=True, t0=0.0):
        if self.vRest is None:
            self.settle()
        if self.eyeV0 is not None:
            for s in range(2):
                self.eyes[s].v = self.eyeV0[s].copy()
        epochs = -(-int(steps) // self.epoch)
        sched = np.zeros((epochs, self.nD), np.float32)
        aAt = np.zeros((epochs, self.nD), np.float32)
        for e in self.eyes:
            if record:
                e.start_tape()
            else:
                e.tape = None
        for ep in range(epochs):
            t = t0 + ep * self.dt_fv
            for s in range(2):
                eye = self.eyes[s]
                eye.setInput(lum(s, t))
                eye.step()
                for (n, node) in self.pairs[s]:
                    a = float(eye.v[node]) - float(self.vRest[node])
                    r = self._rate(a)
                    r32 = np.float32(r)
                    if r > 0 and r32 == 0:          # in the driven set with a zero drive
                        r32 = np.float32(-0.0)
                    sched[ep, self.slot[n]] = r32
                    aAt[ep, self.slot[n]] = a
        self.net.set_drive_schedule(self.driveIdx, self.epoch, sched, drive_soft=self.driveSoft)
        L = self.net.forward(int(steps), seed=seed, record=record)
        self.sched, self.aAt, self.epochs = sched, aAt, epochs
        return L

    def coupling_backward(self, gDrive):
        """src/visdiff.js:211-233, pair by pair."""
        c = self.c
        gDrive = np.asarray(gDrive, np.float32).reshape(self.epochs, self.nD)
        dLdV = [np.zeros((self.epochs, self.fvN), np.float32) for _ in range(2)]
        gGain, live, dcount = 0.0, 0, 0
        for e in range(self.epochs):
            for s in range(2):
                for (n, node) in self.pairs[s]:
                    k = self.slot[n]
                    a = float(self.aAt[e, k])
                    dcount += 1
                    dr = self._drate(a)
                    if dr == 0:
                        continue
                    if a > c["dead"] and c["gain"] * a < c["cap"]:
                        live += 1
                    dR = float(gDrive[e, k])
                    if dR == 0:
                        continue
                    gGain += dR * self._drate_gain(a)
                    dLdV[s][e, node] = np.float32(float(dLdV[s][e, node]) + dR * dr)
        return dLdV, gGain, (live / dcount if dcount else 0.0)

    def backward(self, dLdSpike, truncate=0, lum=False):
        self.net.set_loss(dLdSpike, 0)
        cns = self.net.backward(truncate)
        gDrive = self.net.drive_grad()
        dLdV, gGain, live = self.coupling_backward(gDrive)
        g = {"weight": np.zeros(self.fvE, np.float32), "bias": np.zeros(self.fvN, np.float32),
             "kdt": np.zeros(self.fvN, np.float32)}
        gLum = [np.zeros((self.epochs, self.nCol), np.float32) for _ in range(2)] if lum else [None, None]
        for s in range(2):
            self.eyes[s].backward(dLdV[s], g, gLum[s])
        out = {"params": cns["params"], "logGain": cns["logGain"], "gain": gGain, "fv": g, "live": live,
               "drive": gDrive}
        if lum:
            out["lum"] = gLum
        return out


# ---------------------------------------------------------------- the cord lane ensemble

def cord_lane(tables, params: dict, m: int, c: int, steps: int = 1000, kind: str = "oracle") -> np.ndarray:
    """One run of the cord lane ensemble (cordens.bend) on the lif.c checker:
    the per-step leg-pool counts [steps][6] of member m under condition c.

    The configuration restates Cfg.of, Cfg.comm, Comm.factor, Gain.of and
    Cells.go (cordens.bend:62-160): hemilineage gains on the outgoing sign,
    the commissural factor on those edges' weights, tonic bias on
    interneurons, the command on descending cells, 20 Hz on proprioceptors,
    ablated hemilineages silenced by threshold, the raw stream state
    2654435761 + 7919 m + c; the driven list is the descending cells and the
    proprioceptors in ascending order (Ids.go, :401-416).  The loop restates
    Run.go (:248-273): after each step the pools are tallied from its spike
    list, and the proprioceptors' drive for the next step is
    f32(20 + f32(0.6 * total)), total being the pools' count `pdelay` steps
    back (0 before the first step)."""
    bit = lambda b: (m >> b) & 1 == 1
    f32 = np.float32
    g13a, g13b, g19b = (f32(2.0) if bit(k) else f32(0.5) for k in (1, 2, 3))
    tonic = f32(4.0) if bit(4) else f32(0.0)
    pdelay = 40 if bit(5) else 10
    abl = 1 if c == 1 else 2 if c == 2 else 4 if c == 3 else 0
    cmd = f32(0.0) if c == 5 else f32(60.0)
    grp = 2 if c == 4 else (1 if bit(0) else 0)
    factor = (f32(1.0), f32(3.0), f32(0.0))[grp]
    N = int(tables["indptr"].size - 1)
    hl, role = tables["ens:hl"], tables["ens:role"]
    W = np.array(tables["W"], np.float32)
    comm = tables["ens:comm"] > 0
    W[comm] = W[comm] * factor
    sign = np.array(tables["sign"], np.float32)
    for i in range(N):
        h = int(hl[i])
        g = (g13a if h & 1 else f32(1.0)) * (g13b if h & 2 else f32(1.0))
        g = g * (g19b if h & 4 else f32(1.0))
        sign[i] = sign[i] * g

    class _G:
        pass
    gr = _G()
    gr.N, gr.indptr, gr.indices = N, tables["indptr"], tables["indices"]
    o = OracleBrain(gr, params, seed=1, kind=kind, W=W, sign=sign)
    o.f["set_rng"](o.h, (2654435761 + m * 7919 + c) & 0xFFFFFFFF)
    prop = np.zeros(N, bool)
    prop[tables["ens:proprio"]] = True
    o.array("bias")[:] = np.where(role == 0, tonic, f32(0.0))
    if abl:
        o.array("thr")[:] = np.where((hl & abl) != 0, f32(1e6), f32(0.0))
    drive = o.array("drive")
    drive[:] = 0.0
    drive[role == 2] = cmd
    drive[prop] = f32(20.0)
    o.drivenSet = [int(i) for i in np.flatnonzero((role == 2) | prop)]
    o._driven_dirty = True
    pool = tables["ens:poolId"]
    targets = tables["ens:proprio"]
    totals = [0]                       # totals[k] = the pools' count at step k; step 0 has none
    out = np.zeros((steps, 6), np.uint16)
    for k in range(1, steps + 1):
        fired = o.step()
        cnt = np.bincount(pool[fired], minlength=7)[:6] if len(fired) else np.zeros(6, np.int64)
        out[k - 1] = np.minimum(cnt, 65535)
        totals.append(int(cnt.sum()))
        back = totals[k - pdelay] if k - pdelay >= 1 else 0
        drive[targets] = f32(20.0) + f32(0.6) * f32(back)
    return out
