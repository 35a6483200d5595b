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
