"""GPU parity through the Node binding: a brain driven from JavaScript
(LIFHip over the N-API addon) produces lif.c's raster, spike counts,
membrane state and RNG state, checked against the CPU oracle."""
import importlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import OracleBrain

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not shutil.which("node"), reason="node is absent")]

HERE = os.path.dirname(os.path.abspath(__file__))
synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")


@pytest.mark.parametrize("engine", ["auto", "launch"])
def test_lifhip_from_node_matches_oracle(tmp_path, engine):
    g = synth.tiny_graph(1500, 40, seed=23, p_count=0.3)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.6).astype(np.float32)
    p = P.config_params()
    drv = list(range(0, 1500, 9))
    ops = [["setDrive", drv, 220.0], ["setBias", list(range(0, 300)), 5.5], ["setThr", 7, 3.0],
           ["steps", 300], ["run", 400], ["steps", 60]]
    spec = {"N": g.N, "seed": 5, "engine": engine, "params": p,
            "graph": {"indptr": g.indptr.tolist(), "indices": g.indices.tolist(), "weights": g.weights.tolist(),
                      "sign": sign.tolist(), "counts": g.counts.tolist()},
            "ops": ops}
    f = tmp_path / "case.json"
    f.write_text(json.dumps(spec))
    print(f"node LIFHip parity: engine={engine}", flush=True)
    out = subprocess.run(["node", os.path.join(HERE, "js", "napi_run.js"), str(f)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout)
    o = OracleBrain(g, p, seed=5, W=g.weights, sign=sign)
    o.setDrive(drv, 220.0)
    o.setBias(list(range(0, 300)), 5.5)
    o.setThr(7, 3.0)
    nf1, r1, _ = o.run(300, raster_cap=300 * g.N)
    o.run(400, raster_cap=400 * g.N)
    nf2, r2, _ = o.run(60, raster_cap=60 * g.N)
    assert res["nfired"] == nf1.tolist() + nf2.tolist()
    assert res["raster"] == r1.tolist() + r2.tolist()
    assert res["spikeCount"] == o.spikeCount.tolist()
    assert np.array_equal(np.array(res["v"], np.float32).view(np.uint32), o.array("v").view(np.uint32))
    assert res["rng"] == o.rng
    assert res["drivenSet"] == drv
    assert res["t"] == pytest.approx(760 * p["dt"])


def _node_case(tmp_path, spec):
    f = tmp_path / "case.json"
    f.write_text(json.dumps(spec))
    out = subprocess.run(["node", os.path.join(HERE, "js", "napi_run.js"), str(f)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def _graph_spec(g, sign):
    return {"indptr": g.indptr.tolist(), "indices": g.indices.tolist(), "weights": g.weights.tolist(),
            "sign": sign.tolist(), "counts": g.counts.tolist()}


def test_queued_setters_match_oracle(tmp_path):
    """Setter calls queue up in JS and reach the device as one fb_apply_deltas
    batch before the next step: thousands of thresholds, repeated additions to
    one conductance (their order matters), a driven neuron removed and added
    back (it moves to the end of the driven list)."""
    g = synth.tiny_graph(1500, 40, seed=29, p_count=0.3)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.6).astype(np.float32)
    p = P.config_params()
    rng = np.random.default_rng(29)
    drv = list(range(0, 1500, 7))
    thr = [[int(i), float(v)] for i, v in zip(rng.integers(0, 1500, 3000), rng.normal(0, 2, 3000))]
    adds = [["addG", 11, 0.7, 0.0], ["addG", 11, 1.3e-3, -0.2], ["addG", 12, 2.0, 0.5], ["addG", 11, 0.25, 0.0]]
    ops = [["setDrive", drv, 220.0], ["steps", 40], ["setThrMany", thr], *adds, ["pulse", [5, 5, 6], 1.5],
           ["setDriveOne", 14, 0.0], ["steps", 30], ["setDriveOne", 14, 300.0], ["setBias", [3, 4, 3], 2.5],
           ["steps", 60]]
    spec = {"N": g.N, "seed": 9, "engine": "auto", "params": p, "graph": _graph_spec(g, sign), "ops": ops}
    print("node LIFHip: queued setters", flush=True)
    res = _node_case(tmp_path, spec)
    o = OracleBrain(g, p, seed=9, W=g.weights, sign=sign)
    o.setDrive(drv, 220.0)
    nfs, ras = [], []
    nf, r, _ = o.run(40, raster_cap=40 * g.N); nfs += nf.tolist(); ras += r.tolist()
    for i, v in thr:
        o.setThr(i, v)
    for _, i, e, ii in adds:
        o.addG(i, e, ii)
    o.pulse([5, 5, 6], 1.5)
    o.setDriveOne(14, 0.0)
    nf, r, _ = o.run(30, raster_cap=30 * g.N); nfs += nf.tolist(); ras += r.tolist()
    o.setDriveOne(14, 300.0)
    o.setBias([3, 4, 3], 2.5)
    nf, r, _ = o.run(60, raster_cap=60 * g.N); nfs += nf.tolist(); ras += r.tolist()
    assert res["nfired"] == nfs
    assert res["raster"] == ras
    assert res["spikeCount"] == o.spikeCount.tolist()
    assert np.array_equal(np.array(res["v"], np.float32).view(np.uint32), o.array("v").view(np.uint32))
    assert res["drivenSet"] == [i for i in drv if i != 14] + [14]


def test_live_mode_batches_like_lifgpu(tmp_path):
    """LIFGpu's protocol (src/lifgpu.js:276-346): steps run in batches of 8,
    setter calls made during a batch act after its steps, those made between
    batches before the next; after flush() the state equals lif.c run with the
    setters at those batch boundaries."""
    g = synth.tiny_graph(1500, 40, seed=31, p_count=0.3)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.6).astype(np.float32)
    p = P.config_params()
    drv = list(range(0, 1500, 9))
    ops = [["setDrive", drv, 220.0], ["steps", 5], ["setThr", 7, 3.0], ["addG", 11, 2.0, 0.0], ["steps", 11],
           ["setBias", [20, 21], 4.0], ["steps", 8], ["flush"], ["setThr", 8, -1.0], ["steps", 3], ["flush"]]
    spec = {"N": g.N, "seed": 4, "engine": "auto", "params": p, "graph": _graph_spec(g, sign), "ops": ops,
            "live": True, "batchSteps": 8, "batchMs": 1e9}
    print("node LIFHip: live mode", flush=True)
    res = _node_case(tmp_path, spec)
    o = OracleBrain(g, p, seed=4, W=g.weights, sign=sign)
    o.setDrive(drv, 220.0)
    o.run(8, raster_cap=8 * g.N)            # batch 1: steps 1-8
    o.setThr(7, 3.0)                        # called during batch 1
    o.addG(11, 2.0, 0.0)
    o.run(8, raster_cap=8 * g.N)            # batch 2: steps 9-16 (submitted by the 16th step)
    o.setBias([20, 21], 4.0)                # called between batches: before batch 3
    o.run(8, raster_cap=8 * g.N)            # batch 3: steps 17-24
    o.setThr(8, -1.0)
    o.run(3, raster_cap=3 * g.N)            # batch 4, submitted by flush()
    assert res["spikeCount"] == o.spikeCount.tolist()
    assert np.array_equal(np.array(res["v"], np.float32).view(np.uint32), o.array("v").view(np.uint32))
    assert res["rng"] == o.rng
    assert res["t"] == pytest.approx(27 * p["dt"])


def test_flyvis_diff_from_node_matches_oracle(tmp_path):
    """FlyVisDiff from JavaScript (flyviship.js over the addon's fv* calls):
    two eyes on one model, taped steps with a luminance frame each, the
    reverse sweep into one gradient; activity and every gradient
    bit-identical to the CPU restatement."""
    from test_flyvis import small_model
    import oracle as orc
    m = small_model(seed=21, N=900, nCol=30)
    rng = np.random.default_rng(21)
    T = 12
    lums = [[rng.uniform(0, 1, 30).astype(np.float32) for _ in range(T)] for _ in range(2)]
    dLdV = []
    for _ in range(2):
        d = np.zeros((T, m["N"]), np.float32)
        d[:, rng.choice(m["N"], 80, replace=False)] = rng.normal(0, 1, (T, 80))
        dLdV.append(d)
    spec = {"model": {"N": m["N"], "E": m["E"], "indptr": m["indptr"].tolist(), "target": m["target"].tolist(),
                      "weight": m["weight"].tolist(), "bias": m["bias"].tolist(), "tau": m["tau"].tolist(),
                      "dt": m["dt"], "inputIdx": {k: np.asarray(v).tolist() for k, v in m["inputIdx"].items()}},
            "lums": [[x.tolist() for x in lums[e]] for e in range(2)], "dLdV": [d.reshape(-1).tolist() for d in dLdV]}
    f = tmp_path / "fv.json"
    f.write_text(json.dumps(spec))
    print("node FlyVisDiffHip parity", flush=True)
    out = subprocess.run(["node", os.path.join(HERE, "js", "fv_run.js"), str(f)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout)
    g = {"weight": np.zeros(m["E"], np.float32), "bias": np.zeros(m["N"], np.float32),
         "kdt": np.zeros(m["N"], np.float32)}
    for e in range(2):
        o = orc.FlyVisDiffOracle(m)
        o.start_tape()
        for t in range(T):
            o.setInput(lums[e][t])
            o.step()
        assert np.array_equal(np.array(res["v"][e], np.float32).view(np.uint32), o.v.view(np.uint32)), e
        dl = np.zeros((T, o.nCol), np.float32)
        o.backward(dLdV[e], g, dl)
        assert np.array_equal(np.array(res["lum"][e], np.float32).view(np.uint32), dl.reshape(-1).view(np.uint32)), e
    for k in ("weight", "bias", "kdt"):
        assert np.array_equal(np.array(res[k], np.float32).view(np.uint32), g[k].view(np.uint32)), k
    assert res["tape"] == T


def test_short_arrays_raise_range_errors():
    """Every addon call that copies a fixed number of elements into or out of
    a caller's TypedArray refuses an array one element short with a RangeError
    (the C side would otherwise read or write past its end)."""
    print("node short-array checks", flush=True)
    out = subprocess.run(["node", os.path.join(HERE, "js", "short_arrays.js")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout)
    for k in ("getArrayOk", "fvBackwardOk"):
        assert res.pop(k) == "no error", k
    assert res and all(v == "RangeError" for v in res.values()), res


# ---------------------------------------------------------------- the gradient and the ensemble from Node

def _node_json(tmp_path, spec, script="diff_run.js"):
    f = tmp_path / "case.json"
    f.write_text(json.dumps(spec))
    out = subprocess.run(["node", os.path.join(HERE, "js", script), str(f)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def test_lifdiffhip_from_node_reproduces_grad_check(tmp_path):
    """The reference's gradient check (scripts/grad_check.mjs 800 120, the
    published problem) driven from JavaScript through LIFDiffHip with
    LIFDiff's own constructor inputs (data {N, indptr, indices, weights, nt},
    opts with P0 and the LCG's arrays): the published adjoint table
    (docs/33-differentiable-brain.md:50-62) to 1e-5, and the oracle's
    gradients to 1e-5."""
    import gradcheck as G
    if not G.available():
        pytest.skip("grad-check subgraph fixture absent")
    print("LIFDiffHip from Node: grad_check", flush=True)
    gc = G.GradCheck(published=True)
    z = np.load(G.FIXTURE, allow_pickle=False)
    data = {"N": gc.N, "indptr": gc.indptr.tolist(), "indices": gc.indices.tolist(),
            "weights": gc.raw_counts.astype(int).tolist(), "nt": z["nt"].astype(int).tolist()}
    opts = {**G.P0, "coba": True, "sizeLog": gc.sizeLog.tolist(), "thrMask": gc.thr_mask.tolist(),
            "biasMask": gc.bias_mask.tolist(), "logGain": gc.logGain.tolist()}
    res = _node_json(tmp_path, {"mode": "diff", "data": data, "opts": opts, "steps": G.STEPS, "seed": 1,
                                "dL": gc.dL.tolist()})
    o = gc.oracle()
    Lo = o.forward(G.STEPS, seed=1)
    go = o.backward()
    # grad_check's loss: the targets' summed spike counts (grad_check.mjs:77-78)
    L = float(np.dot(np.array(res["spikeCount"], np.float64), gc.dL.astype(np.float64)))
    assert abs(L - Lo) <= 1e-5 * abs(Lo)
    for k, pub in G.PUBLISHED.items():
        assert abs(res["params"][k] - pub) <= 1e-5 * abs(pub) + 5e-7, (k, res["params"][k], pub)
        assert abs(res["params"][k] - go["params"][k]) <= 1e-5 * abs(go["params"][k]) + 1e-9, k
    lg = np.array(res["logGain"], np.float32)
    for i, pub in G.PUBLISHED_LOGGAIN.items():
        assert abs(float(lg[i]) - pub) <= 1e-5 * abs(pub) + 5e-7
    assert float(np.abs(lg - go["logGain"]).max()) <= 1e-5 * float(np.abs(go["logGain"]).max())


def test_lifdiffhip_from_node_exact32_is_lifc(tmp_path):
    """exact32 from Node: the tape and the membrane state are the reference
    lif.c's, bit for bit, on the exact32 case with the gate, size scaling,
    outScale, inhGain, depression and adaptation on."""
    import exact32cases as xc
    from oracle import available
    print("LIFDiffHip from Node: exact32", flush=True)
    c = xc.Case()
    data = {"N": c.g.N, "indptr": c.g.indptr.tolist(), "indices": c.g.indices.tolist(),
            "weights": c.g.counts.astype(int).tolist(), "nt": np.where(c.g.exc, 1, 2).tolist()}
    opts = {k: v for k, v in c.p.items() if k != "checkpoint"}
    opts.update({"sizeLog": c.sizeLog.tolist(), "outScale": c.outScale.tolist(), "thrOffset": c.thrOffset.tolist(),
                 "biasMask": c.bias_mask.tolist()})
    drive = [[int(i), float(c.drive[i])] for i in c.driven]
    res = _node_json(tmp_path, {"mode": "diff", "data": data, "opts": opts, "drive": drive, "steps": 400,
                                "seed": 3, "dL": [0.0] * c.g.N})
    b = c.lifc(seed=3, kind="ref" if available("ref") else "oracle")
    nf, ras, _ = b.run(400, raster_cap=400 * c.g.N)
    assert np.array_equal(np.diff(res["spikeOff"]), nf) and res["spikeIdx"] == ras.tolist()
    assert np.array_equal(np.array(res["v"], np.float32).view(np.uint32), b.array("v").view(np.uint32))
    assert res["rng"] == b.rng


def test_lifdiffbatchhip_from_node_equals_separate_networks(tmp_path):
    """Three networks with their own gains, drives, losses and windows in one
    LIFDiffBatchHip from Node give exactly what each gives alone."""
    import diffcases as dc
    print("LIFDiffBatchHip from Node", flush=True)
    pr = dc.Problem(over={"soft": False, "tRef": 2.2})
    g = pr.g
    data = {"N": g.N, "indptr": g.indptr.tolist(), "indices": g.indices.tolist(),
            "weights": g.counts.astype(int).tolist(), "nt": np.where(g.exc, 1, 2).tolist()}
    rng = np.random.default_rng(5)
    members = []
    for b in range(3):
        opts = {**{k: v for k, v in pr.p.items() if k != "checkpoint"}, "minSyn": 1,
                "outScale": (pr.inputs["outScale"] * np.exp(rng.normal(0, 0.05, g.N))).astype(np.float32).tolist(),
                "logGain": pr.inputs["logGain"].tolist(), "thrOffset": pr.inputs["thrOffset"].tolist(),
                "sizeLog": pr.inputs["sizeLog"].tolist(), "thrMask": pr.thr_mask.tolist(),
                "biasMask": pr.bias_mask.tolist()}
        drive = [[int(i), 60.0 + 10 * b] for i in pr.driven]
        dL = np.zeros(g.N, np.float32)
        dL[rng.choice(g.N, 40, replace=False)] = 1.0
        members.append({"opts": opts, "drive": drive, "dL": dL.tolist()})
    res = _node_json(tmp_path, {"mode": "diffbatch", "data": data, "members": members, "steps": [90, 120, 57],
                                "seeds": [3, 4, 5]})
    for b in range(3):
        x, y = res["batched"][b], res["alone"][b]
        assert x["loss"] == y["loss"] and x["spikeIdx"] == y["spikeIdx"] and x["spikeOff"] == y["spikeOff"], b
        assert x["logGain"] == y["logGain"] and x["edge"] == y["edge"], b
        for k in P.DIFF_PARAMS:
            if k != "sizeAlpha":
                assert x["params"][k] == y["params"][k], (b, k)


def test_lifbatchhip_from_node_matches_the_oracle(tmp_path):
    """A 16-member ensemble from Node (own seeds, per-type gains, silenced
    neurons, uneven run lengths): every member equals its own lif.c oracle
    run (spike counts, membrane state, stream)."""
    print("LIFBatchHip from Node", flush=True)
    g = synth.tiny_graph(1200, 30, seed=31, p_count=0.3)
    p = P.config_params()
    B, T = 16, 7
    rng = np.random.default_rng(2)
    type_of = rng.integers(0, T, g.N).astype(np.int32)
    gains = np.exp(rng.normal(0, 0.2, (B, T))).astype(np.float32)
    drv = list(range(0, g.N, 25))
    silence = [[b, rng.choice(g.N, 20, replace=False).tolist()] for b in range(B)]
    steps = rng.integers(100, 300, B).tolist()
    spec = {"mode": "batch", "N": g.N, "params": p, "seeds": [b + 1 for b in range(B)], "typeOf": type_of.tolist(),
            "gains": gains.tolist(), "drive": [drv, 150.0], "silence": silence, "steps": steps, "run": max(steps),
            "graph": {"indptr": g.indptr.tolist(), "indices": g.indices.tolist(), "weights": g.weights.tolist(),
                      "sign": g.sign.tolist(), "counts": g.counts.tolist()}}
    res = _node_json(tmp_path, spec)
    pre = np.repeat(np.arange(g.N), np.diff(g.indptr.astype(np.int64)))
    for b in range(B):
        W = (g.counts.astype(np.float64) * gains[b][type_of[pre]].astype(np.float64)).astype(np.float32)
        o = OracleBrain(g, p, seed=b + 1, W=W, sign=g.sign)
        o.setDrive(drv, 150.0)
        for i in silence[b][1]:
            o.setThr(int(i), 1e6)
        o.run(steps[b], raster_cap=steps[b] * g.N)
        m = res["members"][b]
        assert m["spikeCount"] == o.spikeCount.tolist(), b
        assert np.array_equal(np.array(m["v"], np.float32).view(np.uint32), o.array("v").view(np.uint32)), b
        assert m["rng"] == o.rng, b


def test_lifdiffhip_from_node_spike_train(tmp_path):
    """forward(steps, {spikeTrain}) from Node (src/lifdiff.js:219, 243,
    298-307): scripts/vnc_regression.mjs's burst -- a 20 Hz train on neuron 0
    of a 0 -> 1 edge (count 10, wSyn 5) fires neuron 1, and the tape keeps the
    train -- and on a random network with drive the tape, state, stream and
    gradients of the oracle (bit-identical forward, gradients to 1e-5)."""
    import diffcases as dc
    print("LIFDiffHip from Node: spikeTrain", flush=True)
    data = {"N": 2, "indptr": [0, 1, 1], "indices": [1], "weights": [10], "nt": [1, 1]}
    train = [[0] if (t % 5 == 0 and t < 50) else None for t in range(60)]
    res = _node_json(tmp_path, {"mode": "diff", "data": data, "opts": {"wSyn": 5, "minSyn": 1}, "steps": 60,
                                "seed": 1, "dL": [0.0, 1.0], "spikeTrain": train})
    assert res["trainKept"]
    # dt 0.5 ms, tRef 2.2 ms: a forced spike leaves neuron 0 refractory for
    # 2.7 ms, so every second listing (5 steps = 2.5 ms on) is skipped
    assert res["spikeIdx"].count(1) > 0 and res["spikeIdx"].count(0) == 5
    pr = dc.Problem(over={"soft": False, "tRef": 2.2}, seed=31)
    g = pr.g
    data = {"N": g.N, "indptr": g.indptr.tolist(), "indices": g.indices.tolist(),
            "weights": g.counts.astype(int).tolist(), "nt": np.where(g.exc, 1, 2).tolist()}
    opts = {**{k: v for k, v in pr.p.items() if k != "checkpoint"},
            "outScale": pr.inputs["outScale"].tolist(), "logGain": pr.inputs["logGain"].tolist(),
            "thrOffset": pr.inputs["thrOffset"].tolist(), "sizeLog": pr.inputs["sizeLog"].tolist(),
            "thrMask": pr.thr_mask.tolist(), "biasMask": pr.bias_mask.tolist()}
    rng = np.random.default_rng(9)
    train = [rng.integers(0, g.N, size=rng.integers(0, 6)).tolist() if rng.random() < 0.6 else None
             for _ in range(100)]
    drive = [[int(i), 60.0] for i in pr.driven]
    res = _node_json(tmp_path, {"mode": "diff", "data": data, "opts": opts, "drive": drive, "steps": 120, "seed": 3,
                                "dL": pr.dL.tolist(), "spikeTrain": train})
    o = pr.oracle()
    Lo = o.forward(120, seed=3, spikeTrain=train)
    off, idx, _, forced = o.spikes()
    ras = []
    for t in range(120):   # lif.c's order: forced as listed, then threshold spikes by index
        seg, fs = idx[off[t]:off[t + 1]], forced[off[t]:off[t + 1]]
        ras += seg[fs].tolist() + sorted(seg[~fs].tolist())
    assert res["spikeOff"] == off.tolist() and res["spikeIdx"] == ras
    assert np.array_equal(np.array(res["v"], np.float32).view(np.uint32), o.state("v").view(np.uint32))
    # the loss: the targets' summed spike counts (LIFDiff's forward reports
    # none; the caller's dLdSpike goes to backward)
    L = float(np.dot(np.array(res["spikeCount"], np.float64), pr.dL.astype(np.float64)))
    assert res["rng"] == o.rng and abs(L - Lo) <= 1e-12 * abs(Lo)
    go = o.backward()
    for k in P.DIFF_PARAMS:
        assert abs(res["params"][k] - go["params"][k]) <= 1e-5 * abs(go["params"][k]) + 1e-9, k
    lg = np.array(res["logGain"], np.float32)
    assert float(np.abs(lg - go["logGain"]).max()) <= 1e-5 * float(np.abs(go["logGain"]).max())
