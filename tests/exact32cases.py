"""exact32 cases: a LIFDiff network in exact32 mode (src/lifdiff.js:105-109,
173-190, 231-240, 263-274, 317-326) and the lif.c brain it must equal bit for
bit -- the reference's twin audit (scripts/twin_audit.mjs:107-144) on
synthetic graphs with the minSyn gate, size scaling (inScale), outScale,
inhGain, depression, adaptation, a masked bias and per-neuron threshold
offsets switched on.

The lif.c side gets writeGraph's arrays as LIFDiff bakes them (W[k] =
f32(count * inScale[q] * (sign < 0 ? inhGain : 1) * outScale[j]), sign =
f32(EXC_SIGN * wSyn)), bias = laminaBias on the masked neurons, thr =
thrOffset, and the same driven list and seed."""
import importlib

import numpy as np

synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")


class Case:
    def __init__(self, N=600, deg=30.0, seed=5, depU=0.15, adaptInc=0.4, over=None, drive_every=7, hz=80.0):
        g = synth.tiny_graph(N, deg, seed=seed, p_count=0.3)
        self.g = g
        rng = np.random.default_rng(seed + 77)
        self.p = P.diff_merged({"soft": False, "exact32": True, "minSyn": 2, "wSyn": 0.45, "sizeAlpha": 0.5,
                                "inhGain": 0.8, "eInh": -76.0, "laminaBias": 9.0, "adaptInc": adaptInc,
                                "depU": depU, "tRef": 2.2, **(over or {})})
        self.sign = np.where(g.exc, 1.0, -1.0).astype(np.float32)
        self.counts = P.gated_counts(g.counts, g.indices, self.p["minSyn"])
        self.sizeLog = (rng.random(N) * 2 - 1).astype(np.float32)
        self.outScale = (0.5 + rng.random(N)).astype(np.float32)
        self.thrOffset = ((rng.random(N) - 0.5) * 4).astype(np.float32)
        self.bias_mask = (np.arange(N) < N // 4).astype(np.uint8)
        self.driven = np.arange(3, N, drive_every, dtype=np.int32)
        self.drive = np.zeros(N, np.float32)
        self.drive[self.driven] = hz

    def in_scale(self):
        """inScale = min(boostCap, max(1/maxSizeScale, exp(-sizeAlpha sizeLog)))
        in binary64, stored binary32 (src/lifdiff.js:161-164)."""
        s = np.exp(-self.p["sizeAlpha"] * self.sizeLog.astype(np.float64))
        return np.minimum(self.p["boostCap"], np.maximum(1.0 / self.p["maxSizeScale"], s)).astype(np.float32)

    def baked(self):
        """writeGraph's W and sign as exact32 bakes them (src/lifdiff.js:182-189)."""
        g = self.g
        pre = np.repeat(np.arange(g.N), np.diff(g.indptr.astype(np.int64)))
        ig = np.where(self.sign < 0, self.p["inhGain"], 1.0) * self.outScale.astype(np.float64)
        W = (self.counts.astype(np.float64) * self.in_scale().astype(np.float64)[g.indices] * ig[pre]).astype(np.float32)
        return W, (self.sign.astype(np.float64) * self.p["wSyn"]).astype(np.float32)

    def apply(self, net):
        net.set_input("sizeLog", self.sizeLog)
        net.set_input("outScale", self.outScale)
        net.set_input("thrOffset", self.thrOffset)
        net.set_input("drive", self.drive)
        net.set_masks(np.zeros(self.g.N, np.uint8), self.bias_mask)
        net.set_driven(self.driven)
        net.set_loss(np.zeros(self.g.N, np.float32), 0)
        return net

    def oracle(self, params=None):
        from oracle import OracleDiff
        return self.apply(OracleDiff(self.g.N, self.g.indptr, self.g.indices, self.counts, self.sign,
                                     {**self.p, **(params or {})}))

    def lifc(self, seed, kind="ref"):
        """The lif.c brain (the reference's own lif.c when built, else its restatement)."""
        from oracle import OracleBrain
        W, sgn = self.baked()
        p = self.p
        lp = P.merged({k: p[k] for k in ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay",
                                          "adaptInc", "adaptTau", "depU", "depTau", "eExc", "eInh")})
        lp.update({"coba": True, "bgRate": 0.0, "preInh": 0.0})
        b = OracleBrain(self.g, lp, seed=seed, kind=kind, W=W, sign=sgn)
        b.setBias(np.flatnonzero(self.bias_mask).tolist(), p["laminaBias"])
        for i in range(self.g.N):
            b.setThr(i, float(self.thrOffset[i]))
        for i in self.driven.tolist():
            b.setDriveOne(i, float(self.drive[i]))
        return b
