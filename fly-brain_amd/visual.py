"""The optic lobe and the CNS as one differentiable model on the GPU: the
VisualChain of src/visdiff.js over FlyVisDiffHip (two eyes on one model) and
LIFDiffHip (the CNS, driven through its drive schedule).

The chain (src/visdiff.js:11-20):

    luminance per hex column -> flyvis, one step per epoch (dt 20 ms)
      -> rate = coupling(v[node] - vRest[node]) on each (neuron, node) pair
      -> drive of the paired CNS neurons, held for an epoch of LIF steps
      -> spikes, a loss on CNS spike counts

and the adjoint back through all of it: the CNS parameters, the coupling
gain, the optic lobe's weight, bias and kdt, and the luminance.

The optic lobe never reads the CNS, so the rates every epoch start would set
(the reference's onEpoch hook, src/visdiff.js:179-191) are computed first,
eye step by eye step, and handed to the CNS as its drive schedule; the CNS
forward then runs on the GPU without returning to the host.  The coupling's
arithmetic is JavaScript's: binary64 on binary32 inputs, binary32 stores
(the drive and the recorded deviation aAt are Float32Arrays)."""
from __future__ import annotations

import math

import numpy as np

from . import params as _params
from .diff import LIFDiffHip
from .flyvis import FlyVisDiffHip

# src/visdiff.js:50 and :68
COUPLING = {"gain": 250.0, "dead": 0.02, "cap": 200.0}
COUPLING_SOFT = {"gateK": 8000.0, "capKnee": 0.05}
_OWN = ("gain", "dead", "cap", "fvDt", "coupling", "driveSoft", "sensoryMask")


def _srelu(x, e):
    return 0.5 * (x + np.sqrt(x * x + e * e))


def coupling_rate(a, c: dict) -> np.ndarray:
    """_rate (src/visdiff.js:131-138), elementwise in binary64."""
    a = np.asarray(a, dtype=np.float64)
    gain, dead, cap = c["gain"], c["dead"], c["cap"]
    if c.get("coupling") != "soft":
        return np.where(a <= dead, 0.0, np.minimum(cap, gain * a))
    K, knee = COUPLING_SOFT["gateK"], COUPLING_SOFT["capKnee"]
    with np.errstate(over="ignore"):
        gate = 1.0 / (1.0 + np.exp(-K * (a - dead)))
    inner = cap - _srelu(cap - gain * a, knee)
    r = inner * gate
    return np.where(r > 0, r, 0.0)       # Math.max(0, r): +0 for -0 and negatives


def coupling_drate(a, c: dict) -> np.ndarray:
    """_drate (src/visdiff.js:141-149): d rate / d deviation."""
    a = np.asarray(a, dtype=np.float64)
    gain, dead, cap = c["gain"], c["dead"], c["cap"]
    if c.get("coupling") != "soft":
        return np.where((a <= dead) | (gain * a >= cap), 0.0, gain)
    K, knee = COUPLING_SOFT["gateK"], COUPLING_SOFT["capKnee"]
    u = cap - gain * a
    inner = cap - _srelu(u, knee)
    d_inner = 0.5 * (1.0 + u / np.sqrt(u * u + knee * knee))
    with np.errstate(over="ignore"):
        gate = 1.0 / (1.0 + np.exp(-K * (a - dead)))
    return gain * d_inner * gate + inner * gate * (1.0 - gate) * K


def coupling_drate_gain(a, c: dict) -> np.ndarray:
    """_drateGain (src/visdiff.js:151-159): d rate / d gain."""
    a = np.asarray(a, dtype=np.float64)
    gain, dead, cap = c["gain"], c["dead"], c["cap"]
    if c.get("coupling") != "soft":
        return np.where((a <= dead) | (gain * a >= cap), 0.0, a)
    K, knee = COUPLING_SOFT["gateK"], COUPLING_SOFT["capKnee"]
    u = cap - gain * a
    d_inner = 0.5 * (1.0 + u / np.sqrt(u * u + knee * knee))
    with np.errstate(over="ignore"):
        gate = 1.0 / (1.0 + np.exp(-K * (a - dead)))
    return a * d_inner * gate


def _drive_value(rate) -> np.ndarray:
    """The binary32 drive setDriveOne stores for a binary64 rate, with -0.0
    for a positive rate too small for binary32: setDriveOne still puts that
    neuron in the driven set (it tests the binary64 rate), and the drive
    schedule reads -0.0 as exactly that."""
    r = np.asarray(rate, dtype=np.float64)
    r32 = r.astype(np.float32)
    return np.where((r > 0) & (r32 == 0), np.float32(-0.0), r32)


class VisualChainHip:
    """``new VisualChain(data, fvModel, fvMap, opts)`` (src/visdiff.js:70-119).

    ``cns``: {N, indptr, indices, counts, sign} as LIFDiffHip takes them
    (counts before the minSyn / sensory gate); ``fv_model``: parse_flyvis's
    dict (or one of the same form); ``fv_map``: {"eyes": {"L": {"pairs":
    [[neuron, node], ...]}, "R": {...}}}; ``opts``: the CNS options, plus
    gain, dead, cap, fvDt (ms), coupling ("hard" | "soft"), driveSoft and
    sensoryMask."""

    def __init__(self, cns: dict, fv_model: dict, fv_map: dict, opts: dict | None = None):
        opts = dict(opts or {})
        self.c = {**COUPLING, **{k: opts[k] for k in ("gain", "dead", "cap", "coupling") if k in opts}}
        self.sides = ("L", "R")
        self.eyes = [FlyVisDiffHip(fv_model)]
        self.eyes.append(FlyVisDiffHip(shared=self.eyes[0].sharedParts))
        self.fvN, self.fvE, self.nCol = self.eyes[0].N, self.eyes[0].E, self.eyes[0].nCol
        self.pairs = []
        for sd in self.sides:
            p = np.asarray(fv_map["eyes"][sd]["pairs"], dtype=np.int64).reshape(-1, 2)
            self.pairs.append({"neuron": p[:, 0].astype(np.int32), "node": p[:, 1].astype(np.int32)})
        allN = np.concatenate([P["neuron"] for P in self.pairs])
        if np.unique(allN).size != allN.size:
            raise ValueError("a CNS neuron is driven by both eyes: the drive slots would collide")
        self.driveIdx = allN.astype(np.int32)
        self.nD = int(allN.size)
        N = int(cns["N"])
        slot = np.full(N, -1, np.int64)
        slot[self.driveIdx] = np.arange(self.nD)
        self.pairSlot = [slot[P["neuron"]] for P in self.pairs]
        # per eye, the rank of each pair among the pairs that read the same
        # node (the adjoint adds into dLdV in pair order, one node at a time)
        self._rank = []
        for P in self.pairs:
            order = np.argsort(P["node"], kind="stable")
            nd = P["node"][order]
            start = np.r_[0, np.flatnonzero(np.diff(nd)) + 1]
            r = np.arange(nd.size) - np.repeat(start, np.diff(np.r_[start, nd.size]))
            rank = np.empty(nd.size, np.int64)
            rank[order] = r
            self._rank.append(rank)
        # a coupled neuron is driven from outside: its incoming synapses carry
        # nothing (src/visdiff.js:98-110)
        sensory = np.zeros(N, np.uint8)
        if opts.get("sensoryMask") is not None:
            sensory[:] = np.asarray(opts["sensoryMask"], dtype=np.uint8)
        sensory[self.driveIdx] = 1
        self.sensoryMask = sensory
        cns_opts = {k: v for k, v in opts.items() if k not in _OWN}
        p = _params.diff_merged(cns_opts)
        counts = _params.gated_counts(cns["counts"], cns["indices"], p["minSyn"], sensory)
        self.net = LIFDiffHip(N, cns["indptr"], cns["indices"], counts, cns["sign"], cns_opts)
        cnsDt = float(p["dt"])
        fvDt = float(opts["fvDt"]) if opts.get("fvDt") is not None else self.eyes[0].dt * 1000.0
        self.epoch = max(1, int(math.floor(fvDt / cnsDt + 0.5)))      # 40 LIF steps per optic-lobe step
        self.driveSoft = bool(opts.get("driveSoft", False))
        self.vRest = None
        self.eyeV0 = None
        self.aAt = None
        self.epochs = 0

    def settle(self, level: float = 0.5, steps: int = 150) -> np.ndarray:
        """Resting activity under a uniform grey field (src/visdiff.js:121-128)."""
        grey = np.full(self.nCol, level, np.float32)
        for e in self.eyes:
            e.settle(grey, steps)
        self.vRest = self.eyes[0].v
        self.eyeV0 = [e.v for e in self.eyes]
        return self.vRest

    def forward(self, steps: int, seed=1, lum=None, record: bool = True, t0: float = 0.0) -> float:
        """forward (src/visdiff.js:169-197).  ``lum(eye, t_seconds)`` gives the
        luminance per hex column.  Returns the CNS loss of the loss set on
        ``net`` (0 when none is set)."""
        if self.vRest is None:
            self.settle()
        if self.eyeV0 is not None:
            for s in range(2):
                self.eyes[s].set_v(self.eyeV0[s])
        epochs = -(-int(steps) // self.epoch)
        fvDt = self.eyes[0].dt
        sched = np.zeros((epochs, self.nD), np.float32)
        aAt = np.zeros((epochs, self.nD), np.float32) if record else None
        if record:
            for e in self.eyes:
                e.startTape(epochs)
        vRest = self.vRest.astype(np.float64)
        for ep in range(epochs):
            t = t0 + ep * fvDt
            for s in range(2):
                eye = self.eyes[s]
                eye.setInput(lum(s, t))
                if record:
                    eye.stepTaped()
                else:
                    eye.step()
                P, sl = self.pairs[s], self.pairSlot[s]
                a = eye.v[P["node"]].astype(np.float64) - vRest[P["node"]]
                sched[ep, sl] = _drive_value(coupling_rate(a, self.c))
                if record:
                    aAt[ep, sl] = a
        self.net.set_drive_schedule(self.driveIdx, self.epoch, sched, drive_soft=self.driveSoft)
        L = self.net.forward(int(steps), seed=seed, record=record)
        self.sched = sched
        self.aAt = aAt
        self.epochs = epochs
        return L

    def coupling_backward(self, gDrive: np.ndarray):
        """The join's adjoint (src/visdiff.js:211-233): dLdV per eye
        [epochs][fvN] and the coupling gain's gradient from the CNS drive
        gradient [epochs][nD]; also the live fraction."""
        c = self.c
        gDrive = np.asarray(gDrive, dtype=np.float32).reshape(self.epochs, self.nD)
        aAt = self.aAt
        dLdV = [np.zeros((self.epochs, self.fvN), np.float32) for _ in range(2)]
        terms = []
        live = 0
        for e in range(self.epochs):
            for s in range(2):
                sl = self.pairSlot[s]
                a = aAt[e, sl].astype(np.float64)
                dr = coupling_drate(a, c)
                on = dr != 0.0
                live += int(np.count_nonzero(on & (a > c["dead"]) & (c["gain"] * a < c["cap"])))
                dR = gDrive[e, sl].astype(np.float64)
                use = on & (dR != 0.0)
                terms.append(np.where(use, dR * np.where(use, coupling_drate_gain(a, c), 0.0), 0.0))
                x = dR * dr
                node = self.pairs[s]["node"]
                row = dLdV[s][e]
                rank = self._rank[s]
                for r in range(int(rank.max()) + 1 if rank.size else 0):
                    m = use & (rank == r)
                    nd = node[m]
                    row[nd] = (row[nd].astype(np.float64) + x[m]).astype(np.float32)
        t = np.concatenate(terms) if terms else np.zeros(0)
        gGain = float(np.cumsum(t)[-1]) if t.size else 0.0        # a running sum in pair order
        dcount = self.epochs * self.nD
        return dLdV, gGain, (live / dcount if dcount else 0.0)

    def backward(self, dLdSpike, truncate: int = 0, lum: bool = False) -> dict:
        """backward (src/visdiff.js:207-241): {params, logGain, gain, fv:
        {weight, bias, kdt}, lum (per eye, [epochs][nCol]) when asked, live,
        clamped}.  ``dLdSpike`` None keeps the loss already set on ``net``
        (with its lossFrom)."""
        cns = self.net.backward(dLdSpike, truncate=truncate)
        gDrive = self.net.drive_grad()
        dLdV, gGain, live = self.coupling_backward(gDrive)
        g = {"weight": np.zeros(self.fvE, np.float32), "bias": np.zeros(self.fvN, np.float32),
             "kdt": np.zeros(self.fvN, np.float32)}
        gLum = [np.zeros(self.epochs * self.nCol, np.float32) for _ in range(2)] if lum else [None, None]
        for s in range(2):
            self.eyes[s].backward(dLdV[s], g, gLum[s])
        out = {"params": cns["params"], "logGain": cns["logGain"], "gain": gGain, "fv": g, "live": live,
               "clamped": cns["clamped"], "drive": gDrive}
        if lum:
            out["lum"] = [x.reshape(self.epochs, self.nCol) for x in gLum]
        return out

    def stats(self) -> dict:
        """stats (src/visdiff.js:245-257)."""
        c = self.c
        a = self.aAt.astype(np.float64).reshape(-1)
        below = int(np.count_nonzero(a <= c["dead"]))
        sat = int(np.count_nonzero((a > c["dead"]) & (c["gain"] * a >= c["cap"])))
        tot = a.size
        meanRate = float(np.cumsum(coupling_rate(a, c))[-1]) / tot if tot else 0.0
        return {"pairs": self.nD, "epochs": self.epochs, "belowDeadband": below / tot,
                "live": (tot - below - sat) / tot, "saturated": sat / tot, "meanRate": meanRate}

    def destroy(self):
        for e in self.eyes[::-1]:
            e.destroy()
        self.net.destroy()
