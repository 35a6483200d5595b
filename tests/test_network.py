"""LIFNetwork, the reference's JS twin (src/lif.js:16-105).

- The CPU restatement (oracle/lifnet_oracle.c) against golden runs of the
  reference's own LIFNetwork with Math.random seeded by mulberry32
  (tests/golden/lifnet_golden.json, written by make_lifnet_golden.mjs).
- The construction path of the engine (fly-brain_amd/network.py and
  LIFHip.fromNetwork in lifhip.js) against the golden baked weights.
- The model-level claim: lif.c's step (what the engine runs, bit for bit)
  and the JS twin are statistically the same model.  Checked here on the CPU
  between the two restatements; test_gpu_network.py checks the engine.
"""
import importlib
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import OracleBrain, OracleNet  # noqa: E402  (the checker)

GOLDEN = os.path.join(HERE, "golden", "lifnet_golden.json")
network = importlib.import_module("fly-brain_amd.network")
P = importlib.import_module("fly-brain_amd.params")
synth = importlib.import_module("fly-brain_amd.synth")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def case_params(c):
    p = dict(c["params"])
    ex = c.get("extra") or {}
    for k in ("inScale", "outScale", "sensoryMask"):
        if k in ex:
            p[k] = ex[k]
    return p


def apply_op(net, op):
    f, *a = op
    if f in ("setDrive", "setBias", "pulse"):
        getattr(net, f)(list(range(a[0], a[1])), a[2])
    elif f == "reset":
        net.reset()
    else:
        getattr(net, f)(*a)


@pytest.mark.parametrize("name", sorted(golden().keys()))
def test_restatement_reproduces_lifnetwork(name):
    c = golden()[name]
    g = c["graph"]
    ex = c.get("extra") or {}
    net = OracleNet(g["N"], g["indptr"], g["indices"], g["weights"], g["nt"], c["params"],
                    inScale=ex.get("inScale"), outScale=ex.get("outScale"), sensoryMask=ex.get("sensoryMask"),
                    rseed=c["rseed"])
    for op in c["ops"]:
        apply_op(net, op)
    for t in range(c["steps"]):
        if c["mid"] and t == c["mid"]["at"]:
            for op in c["mid"]["ops"]:
                apply_op(net, op)
        assert net.step().tolist() == c["fired"][t], f"fired list differs at step {t}"
    fin = c["final"]
    for k in ("v", "gE", "gI", "refr", "trace", "adapt", "res"):
        assert np.array_equal(net.array(k).view(np.uint32), np.array(fin[k], dtype=np.uint32)), k
    assert np.array_equal(net.spikeCount, np.array(fin["spikeCount"], dtype=np.uint32))
    assert net.nAwake == fin["nAwake"]
    assert net.t == fin["t"]


def test_restatement_fails_a_planted_sign_error():
    """A plausible mistake (GABA read as excitatory) must change the golden run."""
    c = golden()["cuba_recurrent"]
    g = c["graph"]
    wrong = [0, 1, 1, -1, 1, 1, 1, -1]
    net = OracleNet(g["N"], g["indptr"], g["indices"], g["weights"], g["nt"], c["params"], ntSign=wrong,
                    rseed=c["rseed"])
    for op in c["ops"]:
        apply_op(net, op)
    fired = [net.step().tolist() for _ in range(c["steps"])]
    assert fired != c["fired"]


@pytest.mark.parametrize("name", sorted(golden().keys()))
def test_bake_matches_lifnetwork(name):
    c = golden()[name]
    g = c["graph"]
    W, sign = network.bake_network(g["N"], g["indptr"], g["indices"], g["weights"], g["nt"], case_params(c))
    assert np.array_equal(W.view(np.uint32), np.array(c["weights_baked"], dtype=np.uint32))
    p = P.merged(c["params"])
    want = (np.array(network.EXC_SIGN)[np.array(g["nt"])] * p["wSyn"]).astype(np.float32)
    assert np.array_equal(sign, want)


@pytest.mark.skipif(not shutil.which("node") or
                    not os.path.exists(os.path.join(ROOT, "fly-brain_amd", "lib", "flybrain_napi.node")),
                    reason="node or the built addon is absent")
def test_js_bake_matches_lifnetwork():
    out = subprocess.run(["node", os.path.join(HERE, "js", "bake_check.js"), GOLDEN], capture_output=True,
                         text=True, timeout=60, check=True)
    res = json.loads(out.stdout)
    assert res and all(res.values()), res


def test_noise_is_refused_before_any_device_call():
    with pytest.raises(ValueError, match="noise"):
        network.LIFNetHip(4, [0, 1, 2, 3, 4], [1, 2, 3, 0], [1, 1, 1, 1], [1, 1, 1, 1], {"noise": 1.0})


# ------------------------------------------------------------------ statistics

STAT_N = 4096
STAT_CONFIGS = {
    "cuba": {"wSyn": 4.0},
    "coba": {"coba": True, "eInh": -80.0, "wSyn": 6.0},
}


def stat_inputs():
    g = synth.lognormal_graph(STAT_N, 12, 0.8, 0.3, 3)
    nt = np.random.default_rng(4).integers(0, 8, STAT_N).astype(np.uint8)
    counts = g.counts.astype(np.float32)
    driven = np.arange(0, STAT_N, 20)
    return g, nt, counts, driven


def twin_counts(params, seeds, steps):
    """Per-neuron spike counts of the JS twin (restated) for each seed."""
    g, nt, counts, driven = stat_inputs()
    out = []
    for s in seeds:
        net = OracleNet(STAT_N, g.indptr, g.indices, counts, nt, params, rseed=s)
        net.setDrive(driven, 150.0)
        net.run(steps)
        out.append(net.spikeCount.astype(np.float64))
    return np.array(out)


def assert_same_model(a, b, driven):
    """a, b: [seeds][N] spike counts of two runs of the model.  The totals
    (whole network and the recurrent, non-driven part) agree within a few
    percent on average over seeds, and the per-neuron recurrent counts are
    strongly correlated."""
    nd = np.ones(a.shape[1], dtype=bool)
    nd[driven] = False
    ta, tb = a.sum(1).mean(), b.sum(1).mean()
    ra, rb = a[:, nd].sum(1).mean(), b[:, nd].sum(1).mean()
    assert abs(ta - tb) <= 0.03 * max(ta, tb), (ta, tb)
    assert abs(ra - rb) <= 0.08 * max(ra, rb), (ra, rb)
    corr = np.corrcoef(a[:, nd].mean(0), b[:, nd].mean(0))[0, 1]
    assert corr > 0.95, corr


class _G:
    pass


@pytest.mark.parametrize("cfg", sorted(STAT_CONFIGS))
def test_lifc_model_matches_lifnetwork_statistically(cfg):
    params = STAT_CONFIGS[cfg]
    g, nt, counts, driven = stat_inputs()
    W, sign = network.bake_network(STAT_N, g.indptr, g.indices, counts, nt, params)
    steps = 4000
    seeds = (1, 2, 3)
    lifc = []
    for s in seeds:
        gg = _G()
        gg.N, gg.indptr, gg.indices = STAT_N, g.indptr, g.indices
        b = OracleBrain(gg, P.merged(params), seed=s, W=W, sign=sign)
        b.setDrive(list(driven), 150.0)
        b.run(steps)
        lifc.append(np.array(b.spikeCount, dtype=np.float64))
    assert_same_model(np.array(lifc), twin_counts(params, seeds, steps), driven)
