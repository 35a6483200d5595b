"""The real-connectome build path (fly-brain_amd/connectome.py) against the
reference's own published answers: the ventral-cord brain built from the
reference's data files reproduces the spike totals of its kernel gate,
1000 steps with seed 1 (docs/46-cord-ir.md:93-96), on the lif.c oracle.
Skipped where the reference's data are absent."""
import importlib
import os

import numpy as np
import pytest

from oracle import OracleBrain

DATA = "/root/reference/public/data"
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(DATA, "cord_ir.bin")),
                                reason="the reference's data files are absent")
C = importlib.import_module("fly-brain_amd.connectome")

# docs/46-cord-ir.md:93-96
CORD_TOTALS = {None: 139582, "hemilineage:13A": 137276, "hemilineage:13B": 142120, "hemilineage:19B": 134097}


class _G:
    pass


def run_oracle(b, steps=1000):
    g = _G()
    g.N, g.indptr, g.indices = b["N"], b["indptr"], b["indices"]
    o = OracleBrain(g, b["params"], seed=1, W=b["W"], sign=b["sign"])
    for i, r in zip(b["driven"], b["rates"]):
        o.setDriveOne(int(i), float(r))
    for i in b["silenced"]:
        o.setThr(int(i), 1e6)
    nf, _, _ = o.run(steps, raster_cap=1)
    return nf


@pytest.mark.parametrize("ablate", list(CORD_TOTALS))
def test_cord_brain_reproduces_the_reference_totals(ablate):
    b = C.cord_brain(DATA, ablate)
    assert int(run_oracle(b).sum()) == CORD_TOTALS[ablate]


def test_cord_build_details():
    """Structure the reference reports for the cord build (docs/46-cord-ir.md):
    21,843 cells, 1,615,182 entries; calibrated parameters (dt 0.5 ms, COBA,
    minSyn 6, neuromodulation on)."""
    b = C.cord_brain(DATA)
    assert b["N"] == 21843 and b["W"].size == 1615182
    p = b["params"]
    assert p["dt"] == 0.5 and p["coba"] is True and p["minSyn"] == 6 and p["depU"] == 0
    # writeGraph's gate: an edge is cut when its count is below minSyn or its target is sensory
    cord = C.load_cord(DATA)
    cut_by_count = cord["counts"].astype(np.int64) < 6
    assert np.all(b["W"][cut_by_count] == 0)


def test_a_planted_scale_error_misses_the_totals():
    """The totals pin the input scale: a sizeAlpha off by 0.01 changes them."""
    b = C.cord_brain(DATA)
    conn = C.load_connectome(DATA)
    cord = C.load_cord(DATA)
    p = dict(b["params"])
    p["sizeAlpha"] = p["sizeAlpha"] + 0.01
    in_scale, sens, _ = C.brain_scales(conn, conn["size"], p)
    sign = C.modulatory_sign(conn, conn["ntsign"], p)
    o = cord["origIdx"]
    W, s = C.write_graph(cord["N"], cord["indptr"], cord["indices"], cord["counts"],
                         np.asarray(sign)[o], b["params"], in_scale[o], sens[o])
    b2 = dict(b, W=W, sign=s)
    assert int(run_oracle(b2, 300).sum()) != int(run_oracle(b, 300).sum())
