"""LIFNetwork's construction path on the HIP engine.

The reference's JS twin ``LIFNetwork`` (src/lif.js:16-105) takes the raw
connectome (synapse counts per edge, a transmitter class per neuron) and bakes
weights and signs in its constructor (src/lif.js:17-33).  ``bake_network``
restates that bake; ``LIFNetHip`` builds a brain from the same constructor
arguments and offers LIFNetwork's method set.

The step is lif.c's, as with LIFWasm: the reference treats the JS twin and
the WASM kernel as statistically equivalent (the same model, its own
comparison giving 11,363 against 11,261 spikes, docs/06-wasm-kernel.md:29).
What differs from the JS twin:
- draws: lif.c's one xorshift32 stream over the driven list, not
  ``Math.random`` per awake driven neuron;
- arithmetic: binary32 throughout, where the JS evaluates in binary64 and
  rounds on store;
- no awake list: quiet neurons decay towards rest instead of snapping to it;
- the fired list is in lif.c's order (driven first, then ascending);
- ``noise`` (Gaussian membrane noise) is not part of lif.c and is refused.
"""
from __future__ import annotations

import numpy as np

from .lif import HipGraph, LIFHip
from .params import merged

# src/lif.js:5: unknown, ACh, GABA, Glu, DA, 5HT, OA, histamine
EXC_SIGN = (0.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0)

# constructor options of LIFNetwork that are arrays, not scalar parameters
ARRAY_OPTIONS = ("inScale", "outScale", "sensoryMask", "ntSign", "preSign")


def presyn_sign(N: int, nt, params: dict) -> np.ndarray:
    """Sign of each presynaptic neuron: ``preSign[j]`` or ``(ntSign || EXC_SIGN)[nt[j]]``."""
    if params.get("preSign") is not None:
        return np.asarray(params["preSign"], dtype=np.float64).reshape(N)
    table = np.asarray(params.get("ntSign") if params.get("ntSign") is not None else EXC_SIGN, dtype=np.float64)
    return table[np.asarray(nt, dtype=np.int64).reshape(N)]


def bake_network(N: int, indptr, indices, weights, nt, params: dict | None = None):
    """(W f32[E], sign f32[N]) as LIFNetwork's constructor bakes them (src/lif.js:17-33).

    ``W[k] = f32(c * inScale[q] * outScale[j])`` for an edge j -> q with count
    ``c >= minSyn`` whose target is not sensory (or, with presynaptic
    inhibition on, whose source is inhibitory); else 0.  Rows of inhibitory
    sources are then multiplied by ``inhGain`` (a second f32 rounding, as the
    JS does it).  ``sign[j] = f32(s_j * wSyn)``, the per-source factor lif.c
    reads.
    """
    p = merged(params)
    N = int(N)
    ip = np.asarray(indptr, dtype=np.int64).reshape(N + 1)
    ix = np.asarray(indices, dtype=np.int64)
    c = np.asarray(weights, dtype=np.float32).astype(np.float64)
    E = int(ip[-1])
    if ix.size != E or c.size != E:
        raise ValueError("indices / weights do not match indptr")
    s = presyn_sign(N, nt, p)
    row = np.repeat(np.arange(N), np.diff(ip))
    inS = np.ones(E) if p.get("inScale") is None else \
        np.asarray(p["inScale"], dtype=np.float32).astype(np.float64)[ix]
    og = np.ones(E) if p.get("outScale") is None else \
        np.asarray(p["outScale"], dtype=np.float32).astype(np.float64)[row]
    sens = np.zeros(E, dtype=bool) if p.get("sensoryMask") is None else \
        np.asarray(p["sensoryMask"]).astype(bool)[ix]
    keep_sensory = (float(p["preInh"]) > 0) & (s[row] < 0)
    on = (c >= float(p["minSyn"])) & (~sens | keep_sensory)
    W = np.where(on, (c * inS * og).astype(np.float32), np.float32(0))
    ig = float(p["inhGain"])
    if ig != 1:
        neg = s[row] < 0
        W[neg] = (W[neg].astype(np.float64) * ig).astype(np.float32)
    sign = (s * float(p["wSyn"])).astype(np.float32)
    return W.astype(np.float32), sign


class LIFNetHip(LIFHip):
    """``new LIFNetwork(N, indptr, indices, weights, nt, params)`` on the HIP engine.

    Keeps LIFNetwork's fields ``indptr``, ``indices``, ``nt``, ``weights``
    (baked), ``sensory`` and its method set (``step``, ``setDrive``,
    ``setDriveOne``, ``setBias``, ``setThr``, ``addG``, ``pulse``,
    ``reset``, ``setParams``, ``wake``).
    """

    def __init__(self, N: int, indptr, indices, weights, nt, params: dict | None = None, seed=1,
                 engine: str = "auto"):
        p = merged(params)
        if float(p.get("noise") or 0.0) > 0:
            raise ValueError("noise > 0: Gaussian membrane noise is not part of the lif.c model")
        W, sign = bake_network(N, indptr, indices, weights, nt, p)
        graph = HipGraph(N, indptr, indices, W, sign)
        super().__init__(graph, {k: v for k, v in p.items() if k not in ARRAY_OPTIONS}, seed=seed, engine=engine)
        self.indptr = np.asarray(indptr, dtype=np.uint32)
        self.indices = np.asarray(indices, dtype=np.uint32)
        self.nt = np.asarray(nt, dtype=np.uint8)
        self.weights = W
        self.sensory = np.zeros(int(N), dtype=np.uint8) if p.get("sensoryMask") is None else \
            np.asarray(p["sensoryMask"], dtype=np.uint8)
        self._own_graph = graph

    def destroy(self):
        super().destroy()
        g = getattr(self, "_own_graph", None)
        if g is not None:
            g.release()
            self._own_graph = None
