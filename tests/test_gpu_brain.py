"""The engine on the reference's whole calibrated brain (165,122 neurons,
10.5 M entries; COBA, dt 0.5 ms, adaptation on, 3,078 driven neurons), from
the compact fixture tests/golden/make_brain_fixture.py writes: the sugar
trace's total over seeds 1-4 equals the shipped kernel's published one
(public/data/twin_audit.json), and seed 1's per-neuron counts and state equal
the oracle's bit for bit."""
import importlib
import os

import numpy as np
import pytest

from oracle import OracleBrain

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "_brain", "calibrated_brain.npz")
lif = importlib.import_module("fly-brain_amd.lif")
C = importlib.import_module("fly-brain_amd.connectome")

# public/data/twin_audit.json, traces[0]: mean over seeds 1-4 of the total
# spike count, and mean spikes per neuron of three reported types
SUGAR_TOTAL_MEAN = 448344.0
SUGAR_TYPE_MEANS = {"MN9": 20.375, "GNG232": 7.25, "DNge080": 9.375}


class _G:
    pass


@pytest.fixture(scope="module")
def brain():
    if not os.path.exists(FIXTURE):
        pytest.skip("the brain fixture was not built (needs the reference's data files)")
    return C.load_compact(FIXTURE)


def _setup(net, b):
    for i in np.nonzero(b["thr"])[0]:
        net.setThr(int(i), float(b["thr"][i]))
    lam = np.nonzero(b["bias"])[0]
    net.setBias(lam, float(b["bias"][lam[0]]))
    net.setDrive(b["orn"], 6.0)
    for i in b["sugar"]:
        net.setDriveOne(int(i), 100.0)


def run_hip(g, b, seed, engine):
    net = lif.LIFHip(g, b["params"], seed=seed, engine=engine)
    _setup(net, b)
    net.advance(800)
    for i in b["sugar"]:
        net.setDriveOne(int(i), 0.0)
    net.advance(600)
    return net


@pytest.mark.parametrize("engine", ["graph", "launch", "persistent", "auto"])
def test_engine_reproduces_the_sugar_trace(brain, engine):
    print(f"calibrated brain: sugar trace, engine={engine}", flush=True)
    g = lif.HipGraph(brain["N"], brain["indptr"], brain["indices"], brain["W"], brain["sign"])
    counts = []
    for seed in (1, 2, 3, 4):
        net = run_hip(g, brain, seed, engine)
        counts.append(net.spikeCount)
        net.destroy()
    g.release()
    assert sum(int(c.sum()) for c in counts) / 4 == SUGAR_TOTAL_MEAN
    for t, want in SUGAR_TYPE_MEANS.items():
        ix = brain["type:" + t]
        assert round(sum(int(c[ix].sum()) for c in counts) / 4 / ix.size, 3) == want


@pytest.mark.parametrize("engine", ["launch", "persistent"])
def test_seed_one_matches_the_oracle(brain, engine):
    print(f"calibrated brain: seed 1 against the oracle, engine={engine}", flush=True)
    g = lif.HipGraph(brain["N"], brain["indptr"], brain["indices"], brain["W"], brain["sign"])
    net = run_hip(g, brain, 1, engine)
    gg = _G()
    gg.N, gg.indptr, gg.indices = brain["N"], brain["indptr"], brain["indices"]
    o = OracleBrain(gg, brain["params"], seed=1, W=brain["W"], sign=brain["sign"])
    _setup(o, brain)
    o.run(800, raster_cap=1)
    for i in brain["sugar"]:
        o.setDriveOne(int(i), 0.0)
    o.run(600, raster_cap=1)
    assert np.array_equal(net.spikeCount, np.array(o.spikeCount))
    for k in ("v", "gE", "gI", "refr", "trace", "adapt"):
        assert np.array_equal(getattr(net, k).view(np.uint32), np.array(o.array(k)).view(np.uint32)), k
    net.destroy()
    g.release()
