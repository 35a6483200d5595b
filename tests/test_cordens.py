"""The cord lane ensemble (cordens.bend) on the CPU checker: runs restated on
the lif.c oracle (oracle.cord_lane) reproduce the reference's published
readings (public/data/cord_search.json, copied into tests/golden/_cord/ by
tests/golden/make_cord_fixture.py): the meanHz of a run -- the upper median
over the six leg pools of each pool's mean rate from 100 ms on -- is fixed by
the pools' spike counts, so an exact match pins the run's tables, gains,
seeds and loop timing.  Also the run order and configuration of the product
module against the reference's run map (scripts/cord_search.mjs:39-50)."""
import importlib
import json
import os

import numpy as np
import pytest

import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
CORD = os.path.join(HERE, "golden", "_cord")
ens = importlib.import_module("fly-brain_amd.cordens")


def _fixture():
    path = os.path.join(CORD, "cord_brains.npz")
    pub = os.path.join(CORD, "cord_search.json")
    if not (os.path.exists(path) and os.path.exists(pub)):
        pytest.skip("the cord fixture was not built (needs the reference's data files)")
    z = np.load(path, allow_pickle=False)
    if "ens:poolId" not in z.files:
        pytest.skip("the cord fixture predates the ensemble tables")
    with open(pub) as f:
        pub = json.load(f)
    want = {(r["member"], ci): rd["meanHz"] for r in pub["members"] for ci, rd in enumerate(r["reads"])}
    return z, json.loads(str(z["params"])), want


# every condition, every axis on and off, both loop delays, the x3 and the cut commissural groups
SAMPLE = [(0, 0), (0, 5), (1, 1), (2, 2), (3, 3), (4, 4), (8, 0), (16, 5), (32, 1), (33, 4), (47, 3), (63, 0)]


def test_run_map_matches_the_reader():
    runs = ens.run_order()
    assert len(runs) == 384 and len(set(runs)) == 384
    assert runs[:6] == [(0, 0), (0, 1), (0, 2), (0, 3), (0, 5), (2, 0)]
    assert runs[160:162] == [(1, 0), (1, 1)] and runs[320:322] == [(0, 4), (1, 4)]
    for m, c in runs:
        cfg = ens.config(m, c)
        assert cfg["comm"] == (2 if c == 4 else m & 1)
        assert cfg["pdelay"] == (40 if m & 32 else 10)
        assert cfg["seed"] == (2654435761 + m * 7919 + c) % 2 ** 32


@pytest.mark.parametrize("m,c", SAMPLE)
def test_lane_reproduces_published_reading(oracle_built, m, c):
    z, p, want = _fixture()
    counts = orc.cord_lane(z, p, m, c, 1000)
    assert counts.shape == (1000, 6) and counts.sum() > 0
    assert ens.mean_hz(counts) == want[(m, c)]
