"""LIFNetwork's construction path on the GPU (fly-brain_amd/network.py,
LIFHip.fromNetwork): the engine runs lif.c's step on LIFNetwork's baked
inputs bit for bit, and that model matches the reference's JS twin
statistically (the restatement of src/lif.js, pinned by golden runs of the
reference itself in test_network.py)."""
import importlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import OracleBrain
from test_network import STAT_CONFIGS, STAT_N, assert_same_model, stat_inputs, twin_counts

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
network = importlib.import_module("fly-brain_amd.network")
P = importlib.import_module("fly-brain_amd.params")


class _G:
    pass


@pytest.mark.parametrize("cfg", sorted(STAT_CONFIGS))
def test_network_brain_runs_lifc_on_the_baked_inputs(cfg):
    params = STAT_CONFIGS[cfg]
    g, nt, counts, driven = stat_inputs()
    print(f"LIFNetHip vs lif.c restatement: {cfg}", flush=True)
    b = network.LIFNetHip(STAT_N, g.indptr, g.indices, counts, nt, params, seed=9)
    b.setDrive(list(driven), 150.0)
    nf_g, ras_g = b.run_record(600, cap=600 * STAT_N)
    W, sign = network.bake_network(STAT_N, g.indptr, g.indices, counts, nt, params)
    assert np.array_equal(b.weights.view(np.uint32), W.view(np.uint32))
    gg = _G()
    gg.N, gg.indptr, gg.indices = STAT_N, g.indptr, g.indices
    o = OracleBrain(gg, P.merged(params), seed=9, W=W, sign=sign)
    o.setDrive(list(driven), 150.0)
    nf_c, ras_c, _ = o.run(600, raster_cap=600 * STAT_N)
    assert np.array_equal(nf_g, nf_c)
    assert np.array_equal(ras_g, ras_c)
    for k in ("v", "gE", "gI", "refr", "res"):
        assert np.array_equal(getattr(b, k).view(np.uint32), o.array(k).view(np.uint32)), k
    b.destroy()


@pytest.mark.parametrize("cfg", sorted(STAT_CONFIGS))
def test_network_brain_matches_the_js_twin_statistically(cfg):
    params = STAT_CONFIGS[cfg]
    g, nt, counts, driven = stat_inputs()
    steps, seeds = 4000, (1, 2, 3)
    print(f"LIFNetHip vs LIFNetwork statistics: {cfg}", flush=True)
    gpu = []
    for s in seeds:
        b = network.LIFNetHip(STAT_N, g.indptr, g.indices, counts, nt, params, seed=s)
        b.setDrive(list(driven), 150.0)
        b.advance(steps)
        gpu.append(b.spikeCount.astype(np.float64))
        b.destroy()
    assert_same_model(np.array(gpu), twin_counts(params, seeds, steps), driven)


@pytest.mark.skipif(not shutil.which("node"), reason="node is absent")
def test_from_network_in_node_matches_python(tmp_path):
    params = STAT_CONFIGS["cuba"]
    g, nt, counts, driven = stat_inputs()
    spec = {"N": STAT_N, "indptr": g.indptr.tolist(), "indices": g.indices.tolist(), "weights": counts.tolist(),
            "nt": nt.tolist(), "params": params, "seed": 4, "engine": "auto", "driven": driven.tolist(),
            "rate": 150.0, "steps": 500}
    f = tmp_path / "net.json"
    f.write_text(json.dumps(spec))
    print("LIFHip.fromNetwork from node", flush=True)
    out = subprocess.run(["node", os.path.join(HERE, "js", "fromnetwork_run.js"), str(f)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout)
    b = network.LIFNetHip(STAT_N, g.indptr, g.indices, counts, nt, params, seed=4)
    b.setDrive(list(driven), 150.0)
    nf = [len(b.step()) for _ in range(500)]
    assert res["nfired"] == nf
    assert res["spikeCount"] == b.spikeCount.tolist()
    b.destroy()
