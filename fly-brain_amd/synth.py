"""Seeded synthetic connectomes and drive schedules for the LIF engine.

The graphs use the layout the reference's data pipeline produces
(scripts/prep_graph.py:33-55 of Lulzx/fly-brain): CSR by presynaptic neuron,
``indptr`` u32[N+1], ``indices`` u32[E] unique and ascending within each row,
and a synapse count per edge.  Weights and signs are baked the way
``writeGraph`` bakes them (src/lifwasm.js:14-34) with inScale, edgeGain,
outScale and inhGain all 1: ``W[k] = f32(count[k])`` and
``sign[j] = f32(+-1 * wSyn)``.

Every draw comes from ``numpy.random.Generator(PCG64(seed))`` in a fixed
order, so a (config, seed) pair names one graph on every machine running the
same numpy.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

WSYN = 0.275


@dataclasses.dataclass
class Graph:
    N: int
    indptr: np.ndarray      # uint32[N+1]
    indices: np.ndarray     # uint32[E]
    counts: np.ndarray      # int32[E] synapse counts
    exc: np.ndarray         # bool[N], True = excitatory presynaptic neuron
    wsyn: float = WSYN

    @property
    def E(self) -> int:
        return int(self.indices.size)

    @property
    def weights(self) -> np.ndarray:
        """``W = f32(count)``: writeGraph's bake with every gain at 1."""
        return self.counts.astype(np.float32)

    @property
    def sign(self) -> np.ndarray:
        """``sign[j] = f32(s * wSyn)`` with s = +1 / -1 (src/lifwasm.js:22)."""
        s = np.where(self.exc, 1.0, -1.0) * self.wsyn
        return s.astype(np.float32)

    @property
    def outdeg(self) -> np.ndarray:
        return np.diff(self.indptr.astype(np.int64))

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.indptr, self.indices, self.counts, self.exc.astype(np.uint8)):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _unique_rows(rng: np.random.Generator, N: int, deg: np.ndarray) -> np.ndarray:
    """Uniform targets without repetition inside a row, sorted ascending.

    Targets are drawn with replacement, duplicates are removed and the short
    rows are topped up until every row has exactly ``deg[i]`` targets.
    Returns the sorted int64 keys ``row * N + target``.
    """
    deg = deg.astype(np.int64)
    rows = np.repeat(np.arange(N, dtype=np.int64), deg)
    keys = np.unique(rows * N + rng.integers(0, N, size=rows.size, dtype=np.int64))
    while True:
        have = np.bincount(keys // N, minlength=N)
        deficit = deg - have
        if not deficit.any():
            return keys
        r2 = np.repeat(np.arange(N, dtype=np.int64), deficit)
        keys = np.unique(np.concatenate([keys, r2 * N + rng.integers(0, N, size=r2.size, dtype=np.int64)]))


def _assemble(rng: np.random.Generator, N: int, deg: np.ndarray, p_count: float,
              exc_frac: float) -> Graph:
    keys = _unique_rows(rng, N, deg)
    indices = (keys % N).astype(np.uint32)
    indptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(deg, out=indptr[1:])
    if indptr[-1] >= 2**32:
        raise ValueError("edge count exceeds the u32 CSR layout")
    counts = (1 + rng.geometric(p_count, size=indices.size)).astype(np.int32)
    exc = rng.random(N) < exc_frac
    return Graph(N=N, indptr=indptr.astype(np.uint32), indices=indices, counts=counts, exc=exc)


def lognormal_graph(N: int, median_deg: float, sigma: float, p_count: float, seed: int,
                    exc_frac: float = 0.6, max_deg: int | None = None) -> Graph:
    """Out-degree ``clamp(round(LogNormal(ln median, sigma)), 1, max_deg)``."""
    rng = _rng(seed)
    hi = N - 1 if max_deg is None else min(max_deg, N - 1)
    deg = np.clip(np.rint(rng.lognormal(np.log(median_deg), sigma, size=N)), 1, hi).astype(np.int64)
    return _assemble(rng, N, deg, p_count, exc_frac)


def scaled_graph(N: int, median_deg: float, sigma: float, E_target: int, max_deg: int,
                 p_count: float, seed: int, exc_frac: float = 0.6) -> Graph:
    """LogNormal out-degrees rescaled so that sum(deg) is E_target within 1%,
    clipped to [1, max_deg], with the largest row set to exactly max_deg."""
    rng = _rng(seed)
    raw = rng.lognormal(np.log(median_deg), sigma, size=N)
    scale = 1.0
    for _ in range(50):
        deg = np.clip(np.rint(raw * scale), 1, max_deg).astype(np.int64)
        deg[int(np.argmax(raw))] = max_deg
        tot = int(deg.sum())
        if abs(tot - E_target) <= 0.002 * E_target:
            break
        scale *= E_target / tot
    return _assemble(rng, N, deg, p_count, exc_frac)


def tiny_graph(N: int, mean_deg: float, seed: int, p_count: float = 0.3,
               exc_frac: float = 0.6) -> Graph:
    """Small Poisson-degree graphs for parity tests (rows may be empty)."""
    rng = _rng(seed)
    deg = np.minimum(rng.poisson(mean_deg, size=N), N).astype(np.int64)
    return _assemble(rng, N, deg, p_count, exc_frac)


# ------------------------------------------------------------------ configs

def config1_graph() -> Graph:
    """16,384 neurons, lognormal out-degree (median 40, sigma 1.0), counts 1+Geom(0.15)."""
    return lognormal_graph(16384, 40.0, 1.0, 0.15, seed=1)


def config2_graph() -> Graph:
    """165,122 neurons, ~8.0M edges, hub rows up to 20,000, counts 1+Geom(1/12)."""
    return scaled_graph(165122, 30.0, 1.1, 8_000_000, 20000, 1.0 / 12.0, seed=2)


def driven_set(N: int, frac: float, seed: int) -> np.ndarray:
    """A seeded subset of ``round(frac*N)`` neurons in ascending order (the
    drivenSet insertion order the configs use)."""
    n = int(round(frac * N))
    return np.sort(_rng(seed).choice(N, size=n, replace=False)).astype(np.int32)


def drive_rates(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    return _rng(seed).uniform(lo, hi, size=n)


def type_ids(N: int, n_types: int = 500, seed: int = 6) -> np.ndarray:
    return _rng(seed).integers(0, n_types, size=N).astype(np.int32)


def type_gains(b: int, n_types: int = 500, sigma: float = 0.2) -> np.ndarray:
    """Instance b's per-type gains, f32(exp(N(0, sigma^2))), seed 100+b."""
    return np.exp(_rng(100 + b).normal(0.0, sigma, size=n_types)).astype(np.float32)


def silenced(b: int, N: int, max_frac: float = 0.05) -> np.ndarray:
    """Instance b's silenced neurons: a U(0, max_frac) fraction, seed 200+b."""
    rng = _rng(200 + b)
    frac = rng.uniform(0.0, max_frac)
    n = int(round(frac * N))
    return np.sort(rng.choice(N, size=n, replace=False)).astype(np.int32)


def run_length(b: int, lo: int = 2000, hi: int = 10000) -> int:
    return int(_rng(300 + b).integers(lo, hi + 1))
