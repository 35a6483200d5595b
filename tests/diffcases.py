"""Problems for the differentiable model shared by the oracle and GPU tests.

``grad_problem`` restates the test problem of the reference's gradient check
(scripts/grad_check.mjs:51-67): a smooth (soft-threshold) network with a type
gain and a threshold offset on every neuron, size scaling, a masked threshold
term and a masked bias, depression and adaptation on, scored on the spike
counts of a random 30% of the neurons, on seeded synthetic graphs of any
size.  The reference's own problem on its 800-neuron subgraph of the real
connectome, with its published adjoint values, is tests/gradcheck.py."""
import importlib

import numpy as np

synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

# P0 of scripts/grad_check.mjs:65-67
P0 = {"soft": True, "tRef": 0.0, "minSyn": 3, "wSyn": 0.5, "sizeAlpha": 0.6, "inhGain": 0.8,
      "eInh": -76.0, "vThresh": -45.0, "kcThreshold": 2.0, "laminaBias": 14.0, "adaptInc": 0.4,
      "depU": 0.15, "surrogateBeta": 1.5}
# finite-difference steps of scripts/grad_check.mjs:86-87
H = {"wSyn": 2e-3, "sizeAlpha": 2e-3, "inhGain": 2e-3, "eInh": 2e-1, "vThresh": 1e-2,
     "kcThreshold": 1e-2, "laminaBias": 1e-2, "adaptInc": 2e-3, "depU": 2e-3}


def fd_ok(a, fd, h, L):
    """grad_check.mjs:92-94: within 0.1% of the gradient plus the f32 loss
    noise floor 4e-6 |L| / h."""
    tol = 1e-3 * abs(fd) + 4e-6 * abs(L) / h
    return abs(a - fd) <= tol, abs(a - fd) / max(1e-12, abs(fd))


class Problem:
    def __init__(self, N=300, deg=24.0, seed=11, over=None, drive_every=8, drive_hz=60.0):
        g = synth.tiny_graph(N, deg, seed=seed, p_count=0.3)
        self.g = g
        self.p = P.diff_merged({**P0, **(over or {})})
        rng = np.random.default_rng(seed + 1000)
        self.sign = np.where(g.exc, 1.0, -1.0).astype(np.float32)
        self.counts = P.gated_counts(g.counts, g.indices, self.p["minSyn"])
        self.inputs = {
            "sizeLog": (rng.random(N) * 2 - 1).astype(np.float32),
            "logGain": ((rng.random(N) - 0.5) * 0.2).astype(np.float32),
            "outScale": (0.5 + rng.random(N)).astype(np.float32),
            "thrOffset": ((rng.random(N) - 0.5) * 4).astype(np.float32),
        }
        self.thr_mask = (rng.random(N) < 0.2).astype(np.uint8)
        self.bias_mask = (np.arange(N) < N / 4).astype(np.uint8)
        self.target = np.flatnonzero(rng.random(N) < 0.3)
        self.dL = np.zeros(N, np.float32)
        self.dL[self.target] = 1.0
        self.driven = np.arange(0, N, drive_every, dtype=np.int32) if drive_every else np.zeros(0, np.int32)
        drive = np.zeros(N, np.float32)
        drive[self.driven] = drive_hz
        self.inputs["drive"] = drive
        self.edge_gain = None
        self.readout = None
        self.sched = None
        self.dper = None

    def with_schedule(self, nD=20, epoch=8, steps=60, soft=True, seed=5, zeros=False):
        """driveIdx / driveEpoch: nD neurons whose rates are set at every
        epoch start (src/lifdiff.js:135-144); ``soft`` = driveSoft.  With
        ``zeros`` some rates are 0, so neurons leave and re-enter the driven set."""
        rng = np.random.default_rng(seed)
        idx = rng.choice(self.g.N, size=nD, replace=False).astype(np.int32)
        nEp = -(-steps // epoch)
        vals = rng.uniform(20, 150, size=(nEp, nD)).astype(np.float32)
        if zeros:
            vals[rng.random(vals.shape) < 0.3] = 0.0
        self.sched = [idx, epoch, vals, soft]
        return self

    def with_readout(self, R=40, tau_ms=20.0, seed=8):
        rng = np.random.default_rng(seed)
        idx = np.sort(rng.choice(self.g.N, size=R, replace=False)).astype(np.int32)
        tgt = rng.uniform(0, 50, size=R).astype(np.float32)
        self.readout = (idx, tgt, tau_ms)
        return self

    def apply(self, net, params=None, inputs=None, edge_gain=None):
        for k, v in {**self.inputs, **(inputs or {})}.items():
            net.set_input(k, v)
        net.set_masks(self.thr_mask, self.bias_mask)
        eg = edge_gain if edge_gain is not None else self.edge_gain
        if eg is not None:
            net.set_edge_gain(eg)
        net.set_driven(self.driven)
        net.set_loss(self.dL, 0)
        if self.readout is not None:
            net.set_readout(*self.readout)
        if self.sched is not None:
            idx, ep, vals, soft = self.sched
            net.set_drive_schedule(idx, ep, vals, drive_soft=soft)
        if self.dper is not None:
            net.set_loss_per_epoch(*self.dper)
        return net

    def oracle(self, params=None, inputs=None, edge_gain=None):
        from oracle import OracleDiff
        p = {**self.p, **(params or {})}
        net = OracleDiff(self.g.N, self.g.indptr, self.g.indices, self.counts, self.sign, p)
        return self.apply(net, params, inputs, edge_gain)
