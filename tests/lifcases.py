"""Forward parity cases shared by the oracle tests and the GPU parity tests.

A case names a seeded graph, parameters, a brain seed and a script of
operations (steps and the LIFWasm setters, src/lifwasm.js:69-78).  ``play``
drives any object with the LIFWasm method set -- the oracle, the reference
build in oracle/_ref, or the HIP engine -- and returns what it observed, so
the same case reads the same way against every implementation.
"""
from __future__ import annotations

import hashlib
import importlib

import numpy as np

synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

STATE = ("v", "gE", "gI", "refr", "trace", "adapt", "res")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:24]


def _graph(spec: dict):
    kind = spec["kind"]
    if kind == "tiny":
        return synth.tiny_graph(spec["N"], spec["deg"], spec["seed"], spec.get("p_count", 0.3))
    if kind == "c1":
        return synth.config1_graph()
    raise ValueError(kind)


def _weights(g, spec: dict):
    """writeGraph-form W and sign; 'wsyn' and 'gain' rescale them like
    p.wSyn and a uniform outScale would."""
    W = g.counts.astype(np.float64) * spec.get("gain", 1.0)
    s = np.where(g.exc, 1.0, -1.0) * spec.get("wsyn", synth.WSYN)
    return W.astype(np.float32), s.astype(np.float32)


CASES = {
    # small CUBA graph, busy: dense drive plus a depolarising bias so recurrent
    # spikes and same-step collisions on one target are common
    "tiny_cuba_busy": dict(
        graph=dict(kind="tiny", N=256, deg=24, seed=11), wspec=dict(wsyn=0.9),
        params=dict(coba=False), seed=3,
        ops=[("drive", (0, 256, 5), 400.0), ("bias", (0, 256, 3), 6.5), ("run", 1500)]),
    # lif.js DEFAULTS: dt 0.5, adaptation and synaptic depression live, COBA
    "tiny_coba_defaults": dict(
        graph=dict(kind="tiny", N=512, deg=30, seed=12), wspec=dict(wsyn=1.2),
        params=dict(dt=0.5, coba=True, adaptInc=2.0, depU=0.2, eInh=-70.0, _defaults=True), seed=5,
        ops=[("drive", (0, 512, 7), 300.0), ("bias", (1, 512, 4), 5.0), ("run", 1200)]),
    # background events (lif.c:46-48)
    "tiny_background": dict(
        graph=dict(kind="tiny", N=1024, deg=16, seed=13), wspec=dict(wsyn=0.6),
        params=dict(coba=False), seed=7,
        ops=[("background", 40.0, 1.5), ("drive", (0, 1024, 50), 150.0), ("run", 1000)]),
    # presynaptic inhibition on driven neurons (lif.c:38)
    "tiny_preinh": dict(
        graph=dict(kind="tiny", N=256, deg=20, seed=14), wspec=dict(wsyn=0.8),
        params=dict(coba=True, preInh=0.5, eInh=-80.0), seed=9,
        ops=[("drive", (0, 256, 4), 350.0), ("bias", (2, 256, 5), 6.0), ("run", 800)]),
    # the setter surface mid-run, including a driven-set reorder and reset()
    "tiny_setters": dict(
        graph=dict(kind="tiny", N=512, deg=20, seed=15), wspec=dict(wsyn=0.7),
        params=dict(coba=False, adaptInc=1.0, depU=0.1), seed=11,
        ops=[("drive", (0, 512, 9), 250.0), ("run", 300), ("pulse", (3, 512, 11), 4.0),
             ("addG", 17, 2.5, -1.25), ("thr", 5, 3.0), ("thr", 10, 1e6),
             ("driveone", 9, 0.0), ("driveone", 9, 180.0), ("bias", (0, 40, 1), 3.3),
             ("run", 300), ("reset",), ("run", 200)]),
    # config 1 graph, config constants, 2% drive at 150 Hz
    "c1_cuba": dict(
        graph=dict(kind="c1"), wspec={}, params=dict(coba=False), seed=1,
        ops=[("drive_c1",), ("run", 2000)]),
    "c1_coba": dict(
        graph=dict(kind="c1"), wspec={}, params=dict(coba=True), seed=1,
        ops=[("drive_c1",), ("run", 2000)]),
}


def case_params(case: dict) -> dict:
    over = dict(case["params"])
    if over.pop("_defaults", False):
        p = P.merged(over)
    else:
        p = P.config_params(**over)
    return p


def build_inputs(name: str):
    case = CASES[name]
    g = _graph(case["graph"])
    W, s = _weights(g, case["wspec"])
    return case, g, W, s, case_params(case)


def play(brain, ops, N: int):
    """Run ``ops`` on ``brain``; returns per-step nfired and the raster."""
    nfs, rasters = [], []
    for op in ops:
        kind = op[0]
        if kind == "run":
            nf, ras = brain_run(brain, op[1])
            nfs.append(nf)
            rasters.append(ras)
        elif kind == "drive":
            a, b, c = op[1]
            brain.setDrive(list(range(a, b, c)), op[2])
        elif kind == "drive_c1":
            brain.setDrive(synth.driven_set(N, 0.02, 3).tolist(), 150.0)
        elif kind == "driveone":
            brain.setDriveOne(op[1], op[2])
        elif kind == "bias":
            a, b, c = op[1]
            brain.setBias(list(range(a, b, c)), op[2])
        elif kind == "pulse":
            a, b, c = op[1]
            brain.pulse(list(range(a, b, c)), op[2])
        elif kind == "addG":
            brain.addG(op[1], op[2], op[3])
        elif kind == "thr":
            brain.setThr(op[1], op[2])
        elif kind == "background":
            brain.setBackground(op[1], op[2])
        elif kind == "reset":
            brain.reset()
        else:
            raise ValueError(kind)
    return np.concatenate(nfs), np.concatenate(rasters)


def brain_run(brain, steps: int):
    if hasattr(brain, "run_record"):
        return brain.run_record(steps, cap=steps * min(brain.N, 512))
    nf, ras, tot = brain.run(steps, raster_cap=steps * brain.N)
    return nf, ras


def summary(brain, nf, ras) -> dict:
    out = {"nfired_sha": sha(nf.astype(np.int32)), "total": int(nf.sum()),
           "raster_sha": sha(ras.astype(np.int32)), "rng": int(brain.rng)}
    for k in STATE:
        out[k] = sha(np.asarray(getattr(brain, k), dtype=np.float32))
    out["spikeCount"] = sha(np.asarray(brain.spikeCount, dtype=np.uint32))
    return out
