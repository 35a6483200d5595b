"""The engine on the reference's own ventral-cord brain (21,843 cells, 1.6 M
entries, calibrated parameters: COBA, dt 0.5 ms, adaptation on), built from the
reference's data files (tests/golden/make_cord_fixture.py): every engine
reproduces the reference kernel's published spike totals
(docs/46-cord-ir.md:93-96) and the oracle's per-neuron counts bit for bit."""
import importlib
import json
import os

import numpy as np
import pytest

from oracle import OracleBrain
from test_connectome import CORD_TOTALS

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "_cord", "cord_brains.npz")
lif = importlib.import_module("fly-brain_amd.lif")


class _G:
    pass


def _inputs():
    if not os.path.exists(FIXTURE):
        pytest.skip("the cord fixture was not built (needs the reference's data files)")
    z = np.load(FIXTURE, allow_pickle=False)
    return z, json.loads(str(z["params"]))


@pytest.mark.parametrize("engine", ["persistent", "graph"])
@pytest.mark.parametrize("ablate", list(CORD_TOTALS))
def test_engine_reproduces_the_cord_totals(engine, ablate):
    z, p = _inputs()
    N = int(z["indptr"].size - 1)
    silenced = z["silenced:" + ablate] if ablate else np.zeros(0, np.int32)
    print(f"cord brain: engine={engine} ablate={ablate}", flush=True)
    g = lif.HipGraph(N, z["indptr"], z["indices"], z["W"], z["sign"])
    b = lif.LIFHip(g, p, seed=1, engine=engine)
    for i, r in zip(z["driven"], z["rates"]):
        b.setDriveOne(int(i), float(r))
    for i in silenced:
        b.setThr(int(i), 1e6)
    b.advance(1000)
    sc = b.spikeCount
    assert int(sc.sum()) == CORD_TOTALS[ablate]
    gg = _G()
    gg.N, gg.indptr, gg.indices = N, z["indptr"], z["indices"]
    o = OracleBrain(gg, p, seed=1, W=z["W"], sign=z["sign"])
    for i, r in zip(z["driven"], z["rates"]):
        o.setDriveOne(int(i), float(r))
    for i in silenced:
        o.setThr(int(i), 1e6)
    o.run(1000, raster_cap=1)
    assert np.array_equal(sc, np.array(o.spikeCount))
    for k in ("v", "gE", "gI", "refr", "adapt"):
        assert np.array_equal(getattr(b, k).view(np.uint32), np.array(o.array(k)).view(np.uint32)), k
    b.destroy()
    g.release()
