"""Pins the CPU oracle (oracle/lif_oracle.c) before anything is checked against it.

1. Golden vectors recorded from the reference's own lif.c (oracle/_ref,
   tests/golden/make_golden.py): per-step spike counts, rasters, final state.
2. Where oracle/_ref is present, a direct run against it.
3. Independent facts: Marsaglia's published xorshift32 value, the GF(2) period
   of the (13,17,5) generator, the seed hash, Math.round semantics of nslots,
   closed forms for the membrane, decay and refractory clock, and brute force
   on 2-3 neuron graphs for delivery delay and in-order accumulation.
"""
import json
import math
import os

import numpy as np
import pytest

import lifcases
import oracle as orc
from oracle import OracleBrain

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "lif_golden.json")))


@pytest.mark.parametrize("name", sorted(lifcases.CASES))
def test_oracle_matches_reference_golden(oracle_built, name):
    case, g, W, s, p = lifcases.build_inputs(name)
    gold = GOLDEN[name]
    assert g.digest() == gold["graph_digest"], "synthetic generator drifted"
    assert lifcases.sha(W) == gold["W_sha"] and lifcases.sha(s) == gold["sign_sha"]
    b = OracleBrain(g, p, seed=case["seed"], kind="oracle", W=W, sign=s)
    nf, ras = lifcases.play(b, case["ops"], g.N)
    assert nf[:64].tolist() == gold["nfired_head"]
    got = lifcases.summary(b, nf, ras)
    for k in ("total", "nfired_sha", "raster_sha", "rng", "spikeCount", *lifcases.STATE):
        assert got[k] == gold[k], k


@pytest.mark.skipif(not orc.available("ref"), reason="oracle/_ref (reference lif.c) not built here")
def test_oracle_matches_reference_build_live(oracle_built):
    case, g, W, s, p = lifcases.build_inputs("tiny_coba_defaults")
    out = []
    for kind in ("oracle", "ref"):
        b = OracleBrain(g, p, seed=case["seed"], kind=kind, W=W, sign=s)
        nf, ras = lifcases.play(b, case["ops"], g.N)
        out.append((nf, ras, b.state_digest()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for k in out[0][2]:
        assert np.array_equal(out[0][2][k].view(np.uint32), out[1][2][k].view(np.uint32)), k


# ---------------------------------------------------------------- RNG

def test_xorshift_marsaglia_value(oracle_built):
    # Marsaglia, "Xorshift RNGs" (J. Stat. Softw. 2003): from y = 2463534242 the
    # (13, 17, 5) xorshift32 returns 723471715 first.
    assert orc.xorshift_stream(2463534242, 1)[0] == 723471715


def _gf2_matrix():
    """Columns of the 32x32 GF(2) matrix of one xorshift32 step, built from
    the generator's definition as a linear map (shift-xor compositions)."""
    def step(x):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        return x
    return [step(1 << b) for b in range(32)]


def _apply(cols, x):
    r = 0
    for b in range(32):
        if (x >> b) & 1:
            r ^= cols[b]
    return r


def _compose(a, b):  # (a o b)
    return [_apply(a, c) for c in b]


def test_xorshift_full_period_and_stream(oracle_built):
    cols = _gf2_matrix()
    # period 2^32 - 1: M^(2^32) == M and M^((2^32-1)/p) != I for each prime p | 2^32-1
    ident = [1 << b for b in range(32)]

    def power(e):
        r, base = ident, cols
        while e:
            if e & 1:
                r = _compose(base, r)
            base = _compose(base, base)
            e >>= 1
        return r
    assert power(2**32 - 1) == ident
    for prime in (3, 5, 17, 257, 65537):
        assert power((2**32 - 1) // prime) != ident
    # the oracle's stream equals repeated application of the matrix
    st = 0x9E3779B9
    stream = orc.xorshift_stream(st, 2000)
    x = st
    for v in stream:
        x = _apply(cols, x)
        assert x == int(v)


def test_seed_hash():
    # (seed * 2654435761) >>> 0 || 1  (src/lifwasm.js:52)
    for seed in (1, 2, 3, 17, 1000, 3_000_000):
        assert orc.seed_hash(seed) == (seed * 2654435761) % 2**32
    assert orc.seed_hash(0) == 1
    assert orc.seed_hash(1) == 2654435761


def test_nslots_math_round():
    lib = orc._lib("oracle")["_lib"]
    assert lib.fbo_nslots(1.8, 0.1) == 19
    assert lib.fbo_nslots(1.8, 0.5) == 5
    # values checked with node's Math.round: 0.35/0.1 = 3.4999999999999996 -> 3,
    # 0.45/0.1 = 4.5 -> 5 (ties toward +infinity), 0.5/1.0 = 0.5 -> 1
    assert lib.fbo_nslots(0.35, 0.1) == 4
    assert lib.fbo_nslots(0.45, 0.1) == 6
    assert lib.fbo_nslots(0.5, 1.0) == 2
    assert lib.fbo_nslots(0.01, 1.0) == 2   # max(1, 0) + 1


# ---------------------------------------------------------------- closed forms

def _lonely(N=2, **over):
    import importlib
    synth = importlib.import_module("fly-brain_amd.synth")
    P = importlib.import_module("fly-brain_amd.params")
    g = synth.Graph(N=N, indptr=np.zeros(N + 1, np.uint32), indices=np.zeros(0, np.uint32),
                    counts=np.zeros(0, np.int32), exc=np.ones(N, bool))
    return g, P.config_params(**over)


def test_membrane_relaxes_to_bias_closed_form(oracle_built):
    g, p = _lonely()
    b = OracleBrain(g, p, seed=1)
    bias = 6.0   # below threshold distance (7 mV): no spike, v -> vRest + bias
    b.setBias([0], bias)
    kM = p["dt"] / p["tauM"]
    vs = []
    for _ in range(3000):
        b.step()
        vs.append(float(b.v[0]))
    n = np.arange(1, 3001)
    closed = p["vRest"] + bias * (1 - (1 - kM) ** n)
    # f32 rounding accumulates to a few 1e-4 mV over 3000 steps; a wrong term is mV-sized
    assert np.max(np.abs(np.array(vs) - closed)) < 1e-3
    assert float(b.v[1]) == np.float32(p["vRest"])   # rest is an exact fixed point


def test_spike_time_and_refractory_clock(oracle_built):
    g, p = _lonely()
    b = OracleBrain(g, p, seed=1)
    bias = 9.0
    b.setBias([0], bias)
    kM = p["dt"] / p["tauM"]
    thr_step = math.ceil(math.log(1 - 7.0 / bias) / math.log(1 - kM))  # closed-form crossing
    fired = [t for t in range(1000) if 0 in b.step().tolist()]
    assert abs(fired[0] + 1 - thr_step) <= 1
    # refractory clock: refr = f32(2.2) decremented by f32(0.1) while > 0
    r, n = np.float32(p["tRef"]), 0
    while r > 0:
        r = np.float32(r - np.float32(p["dt"]))
        n += 1
    # the neuron is clamped for n steps, then integrates from vReset again
    assert fired[1] - fired[0] == n + fired[0] + 1


def test_conductance_decay_stepwise(oracle_built):
    g, p = _lonely()
    b = OracleBrain(g, p, seed=1)
    b.addG(0, 3.0, -2.0)
    dE = np.float32(math.exp(-p["dt"] / p["tauSyn"]))
    ge, gi = np.float32(3.0), np.float32(-2.0)
    for _ in range(500):
        b.step()
        ge, gi = np.float32(ge * dE), np.float32(gi * dE)
        assert b.gE[0] == ge and b.gI[0] == gi


def test_delivery_delay_and_list_order(oracle_built):
    import importlib
    synth = importlib.import_module("fly-brain_amd.synth")
    P = importlib.import_module("fly-brain_amd.params")
    # 0 -> 2 (count 3), 1 -> 2 (count 7); both forced on every non-refractory step
    g = synth.Graph(N=3, indptr=np.array([0, 1, 2, 2], np.uint32), indices=np.array([2, 2], np.uint32),
                    counts=np.array([3, 7], np.int32), exc=np.array([True, True, True]))
    p = P.config_params()
    b = OracleBrain(g, p, seed=1)
    b.setDriveOne(1, 1e7)   # p = drive*dtS = 1000: always fires when not refractory
    b.setDriveOne(0, 1e7)   # driven list order: [1, 0]
    b.addG(2, 0.3, 0.0)     # a non-zero start makes the add order observable
    D = b.nslots - 1
    s = np.float32(0.275)
    a1, a0 = np.float32(np.float32(7) * s), np.float32(np.float32(3) * s)
    dE = np.float32(math.exp(-p["dt"] / p["tauSyn"]))
    ge = np.float32(0.3)
    for t in range(D + 1):
        fired = b.step()
        if t == 0:
            assert fired.tolist() == [1, 0]
        if t < D:
            ge = np.float32(ge * dE)
            assert b.gE[2] == ge
    # arrival at step D: list order 1 then 0, then the decay
    expect = np.float32(np.float32(np.float32(ge + a1) + a0) * dE)
    assert b.gE[2] == expect
