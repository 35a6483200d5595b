"""The cord lane ensemble (cordens.bend, docs/46-cord-ir.md:149-170) on the
GPU: six binary axes over the gains the connectome does not fix, crossed with
six perturbations -- 64 members, 384 runs of 1,000 steps of the reference's
ventral-cord brain (21,843 cells) with the proprioceptive loop closed.

A run is the lif.c brain (the lane kernel, kernel.bend, is lif.c's step with
depression, the trace and the zero-sign test dropped, each an exact no-op for
this brain) configured from its member and condition index alone
(cordens.bend:62-200):

- a hemilineage gain multiplies the cell's outgoing sign, sign' = f32(sign *
  f32(f32(g13A * g13B) * g19B)) over the hemilineages the cell belongs to;
- the commissural edges' weights are multiplied by the group's factor (x1,
  x3, or x0 when the condition cuts them), W' = f32(W * factor);
- tonic drive is a bias on the cord's interneurons (role 0);
- the descending cells are driven at the command rate (0 when the condition
  silences it), the proprioceptors at 20 Hz plus the loop, all in one driven
  list in ascending cell order that no rate change alters;
- an ablated hemilineage never fires by threshold (the lane's alive flag,
  here threshold 1e6);
- the lane's stream starts at the raw state 2654435761 + 7919 m + c;
- after every step the six leg pools' spikes are tallied, and the
  proprioceptors' drive for the next step is 20 + 0.6 * (the pools' total
  `pdelay` steps back), the loop running on the device (fb_set_pool_feedback).

Each distinct (commissural factor, hemilineage gains) pair gets one device
graph, and each (graph, loop delay) pair a set of brains that later runs of
that kind reuse; runs go several at a time, each brain on its own stream."""
from __future__ import annotations

import numpy as np

from . import batch as _batch
from . import lib as _l
from . import lif

MEMBERS, CONDS, STEPS = 64, 6, 1000
CONDITIONS = ("baseline", "ablate_13A", "ablate_13B", "ablate_19B", "no_commissural", "no_command")
AXES = ("commissural x3", "13A gain x2", "13B gain x2", "19B gain x2", "tonic 4 mV", "proprio delay 20 ms")
PROPRIO_HZ, COMMAND_HZ, PGAIN = np.float32(20.0), np.float32(60.0), np.float32(0.6)
N_POOLS = 6


def config(m: int, c: int) -> dict:
    """Cfg.of and Cfg.comm (cordens.bend:76-130): the run's configuration from
    its member and condition."""
    bit = lambda b: (m >> b) & 1 == 1
    return {"g13a": np.float32(2.0 if bit(1) else 0.5), "g13b": np.float32(2.0 if bit(2) else 0.5),
            "g19b": np.float32(2.0 if bit(3) else 0.5), "tonic": np.float32(4.0 if bit(4) else 0.0),
            "pdelay": 40 if bit(5) else 10, "abl": {1: 1, 2: 2, 3: 4}.get(c, 0),
            "cmd": np.float32(0.0) if c == 5 else COMMAND_HZ, "comm": 2 if c == 4 else (1 if bit(0) else 0),
            "seed": (2654435761 + m * 7919 + c) & 0xFFFFFFFF}


COMM_FACTOR = (np.float32(1.0), np.float32(3.0), np.float32(0.0))


def run_order() -> list:
    """The runs in the order cordens.bend writes them: commissural group 0,
    1, 2; within a group, run index ascending (Run.member / Run.cond)."""
    out = []
    for grp in (0, 1, 2):
        n = MEMBERS if grp == 2 else MEMBERS // 2 * 5
        for j in range(n):
            m = j if grp == 2 else 2 * (j // 5) + (grp & 1)
            c = 4 if grp == 2 else (j % 5 if j % 5 < 4 else 5)
            out.append((m, c))
    return out


def hemilineage_gain(hl: np.ndarray, cfg: dict) -> np.ndarray:
    """Gain.of (cordens.bend:147-151), per cell, binary32."""
    one = np.float32(1.0)
    a = np.where(hl & 1, cfg["g13a"], one).astype(np.float32)
    b = np.where(hl & 2, cfg["g13b"], one).astype(np.float32)
    c = np.where(hl & 4, cfg["g19b"], one).astype(np.float32)
    return (a * b) * c


def mean_hz(counts: np.ndarray) -> float:
    """poolMetrics' meanHz (src/exp/pools.js:49-58, 95) of a run read the way
    scripts/cord_search.mjs reads it: 2 ms bins (4 steps) as rates in Hz, from
    100 ms on, each pool's mean, and the upper median over the six pools."""
    n = counts.shape[0] // 4
    rate = (counts[: n * 4].reshape(n, 4, N_POOLS).sum(axis=1).astype(np.float32) * np.float32(500.0))
    i0 = int(np.floor(100.0 / 2.0 + 0.5))
    x = rate[i0:].astype(np.float64)
    means = x.sum(axis=0) / x.shape[0] if x.shape[0] else np.zeros(N_POOLS)
    return float(np.sort(means)[N_POOLS >> 1])


def _batch_max() -> int:
    return 64


class CordEnsemble:
    """``tables``: the cord brain (indptr, indices, W, sign, driven, params)
    with the ensemble's tables (ens:poolId, ens:proprio, ens:hl, ens:role,
    ens:comm), as tests/golden/make_cord_fixture.py writes them."""

    def __init__(self, tables, params: dict, engine: str = "auto"):
        t = tables
        self.p = dict(params)
        self.N = int(t["indptr"].size - 1)
        self.indptr, self.indices = t["indptr"], t["indices"]
        self.W = np.asarray(t["W"], np.float32)
        self.sign = np.asarray(t["sign"], np.float32)
        self.pool = np.asarray(t["ens:poolId"], np.uint8)
        self.proprio = np.asarray(t["ens:proprio"], np.int32)
        self.hl = np.asarray(t["ens:hl"], np.uint8)
        self.role = np.asarray(t["ens:role"], np.uint8)
        self.comm = np.asarray(t["ens:comm"], np.uint8) > 0
        is_prop = np.zeros(self.N, bool)
        is_prop[self.proprio] = True
        self.is_prop = is_prop
        # Ids.go (cordens.bend:401-416): descending cells and proprioceptors, ascending
        self.driven = np.flatnonzero((self.role == 2) | is_prop).astype(np.int32)
        self.engine = engine
        self._graphs = {}
        self._free = {}

    def graph(self, cfg: dict) -> "lif.HipGraph":
        key = (cfg["comm"], float(cfg["g13a"]), float(cfg["g13b"]), float(cfg["g19b"]))
        g = self._graphs.get(key)
        if g is None:
            f = COMM_FACTOR[cfg["comm"]]
            W = np.where(self.comm, self.W * f, self.W).astype(np.float32)
            sign = (self.sign * hemilineage_gain(self.hl, cfg)).astype(np.float32)
            g = lif.HipGraph(self.N, self.indptr, self.indices, W, sign)
            self._graphs[key] = g
        return g

    def _brain(self, cfg: dict, steps: int) -> "lif.LIFHip":
        """A brain for this configuration's graph and loop delay: a finished
        one of the same kind when there is one (its captured step graphs stay
        valid: a run changes data, not kernel arguments), else a new one."""
        key = (cfg["comm"], float(cfg["g13a"]), float(cfg["g13b"]), float(cfg["g19b"]), cfg["pdelay"])
        free = self._free.setdefault(key, [])
        while free:
            b = free.pop()
            if b._cap >= steps:
                b.reset()                                   # state and the loop's history
                return b
            b.destroy()
        b = lif.LIFHip(self.graph(cfg), self.p, seed=1, engine=self.engine)
        _l.check(b.lib.fb_set_drive(b.h, _l.ptr(self.driven), int(self.driven.size), 1.0))   # the list, ascending
        b.set_pool_feedback(self.pool, self.proprio, float(PROPRIO_HZ), float(PGAIN), cfg["pdelay"], steps,
                            n_pools=N_POOLS)
        b._cap, b._key = steps, key
        return b

    def start(self, m: int, c: int, steps: int = STEPS) -> "lif.LIFHip":
        """A run set up and its steps queued on its brain's stream."""
        cfg = config(m, c)
        b = self._brain(cfg, steps)
        b.set_rng(cfg["seed"])
        b.set_array("bias", np.where(self.role == 0, cfg["tonic"], np.float32(0.0)).astype(np.float32))
        thr = np.zeros(self.N, np.float32)
        if cfg["abl"]:
            thr[(self.hl & cfg["abl"]) != 0] = np.float32(1e6)
        b.set_array("thr", thr)
        drive = np.zeros(self.N, np.float32)
        drive[self.role == 2] = cfg["cmd"]
        drive[self.is_prop] = PROPRIO_HZ
        b.set_array("drive", drive)                        # rates; the driven list stays
        b.run_async(steps)
        return b

    def run(self, runs=None, steps: int = STEPS, concurrency: int = 16) -> dict:
        """(m, c) -> per-step pool counts [steps][6] for each run (all 384 by
        default), ``concurrency`` runs in flight, each on its brain's stream."""
        runs = run_order() if runs is None else list(runs)
        out = {}
        for a in range(0, len(runs), concurrency):
            live = [(mc, self.start(*mc, steps=steps)) for mc in runs[a:a + concurrency]]
            for mc, b in live:
                out[mc] = b.pool_counts()[:steps]
                self._free[b._key].append(b)
        return out

    def run_lanes(self, runs=None, steps: int = STEPS) -> dict:
        """The same runs as lanes of the lane ensemble engine, up to 64 to a
        launch (the lanes of kernel.bend): one ensemble per commissural group
        and 64 runs, the hemilineage gains as per-lane type gains on the sign
        (the cell's hemilineage bits are its type), per-lane rates, seeds,
        biases, thresholds and loop delays; the loop runs on the device for
        every lane at once."""
        runs = run_order() if runs is None else list(runs)
        groups = {}
        for mc in runs:
            groups.setdefault(config(*mc)["comm"], []).append(mc)
        interneurons = np.flatnonzero(self.role == 0).astype(np.int32)
        types = np.arange(8, dtype=np.uint8)
        out = {}
        for grp, lst in groups.items():
            f = COMM_FACTOR[grp]
            W = np.where(self.comm, self.W * f, self.W).astype(np.float32)
            g = lif.HipGraph(self.N, self.indptr, self.indices, W, self.sign,
                             np.ones(self.indices.size, np.float32))     # counts: unused in lane mode
            for a in range(0, len(lst), _batch_max()):
                chunk = lst[a:a + _batch_max()]
                cfgs = [config(*mc) for mc in chunk]
                B = len(chunk)
                gains = np.stack([hemilineage_gain(types, c) for c in cfgs]).astype(np.float32)
                bt = _batch.LIFBatch(g, self.p, seeds=np.arange(1, B + 1, dtype=np.float64),
                                     type_of=self.hl.astype(np.int32), gains=gains)
                _l.check(bt.lib.fb_batch_set_drive(bt.h, _l.ptr(self.driven), int(self.driven.size), 1.0))
                lane_drive = np.zeros((B, self.N), np.float32)
                for q, c in enumerate(cfgs):
                    lane_drive[q, self.role == 2] = c["cmd"]
                    lane_drive[q, self.is_prop] = PROPRIO_HZ
                bt.set_lanes(True, lane_drive)
                bt.set_rng(np.array([c["seed"] for c in cfgs], np.uint32))
                for q, c in enumerate(cfgs):
                    if c["tonic"]:
                        bt.setBias(q, interneurons, float(c["tonic"]))
                    if c["abl"]:
                        bt.setThr(q, np.flatnonzero((self.hl & c["abl"]) != 0).astype(np.int32), 1e6)
                bt.set_pool_feedback(self.pool, N_POOLS, self.proprio, float(PROPRIO_HZ), float(PGAIN),
                                     [c["pdelay"] for c in cfgs], steps)
                bt.run(steps)
                rec = bt.pool_counts()
                for q, mc in enumerate(chunk):
                    out[mc] = rec[q, :steps]
                bt.destroy()
            g.release()
        return out

    def release(self):
        for free in self._free.values():
            for b in free:
                b.destroy()
        self._free = {}
        for g in self._graphs.values():
            g.release()
        self._graphs = {}
