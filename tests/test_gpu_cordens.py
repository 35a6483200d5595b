"""The cord lane ensemble on the GPU against the checker and the reference's
published readings: as one brain per run (CordEnsemble.run over
fb_set_pool_feedback, fb_set_rng and the per-step engines) and as lanes of
the lane ensemble engine (CordEnsemble.run_lanes over fb_batch_set_lanes,
fb_batch_set_rng and fb_batch_set_pool_feedback): per-step leg-pool counts
bit-identical to the lif.c checker's runs, and all 384 runs' meanHz equal to
public/data/cord_search.json (copied into tests/golden/_cord/)."""
import importlib

import numpy as np
import pytest

import oracle as orc
from test_cordens import SAMPLE, _fixture

pytestmark = pytest.mark.gpu
ens = importlib.import_module("fly-brain_amd.cordens")


@pytest.mark.parametrize("engine", ["graph", "launch"])
def test_runs_match_checker(engine):
    z, p, want = _fixture()
    e = ens.CordEnsemble(z, p, engine=engine)
    runs = SAMPLE if engine == "graph" else SAMPLE[:4]
    print(f"cord ensemble: {len(runs)} runs, engine {engine}", flush=True)
    got = e.run(runs)
    for m, c in runs:
        ref = orc.cord_lane(z, p, m, c, 1000)
        assert np.array_equal(got[(m, c)], ref), (m, c)
        assert ens.mean_hz(got[(m, c)]) == want[(m, c)]
    e.release()


def test_all_runs_reproduce_published_readings():
    z, p, want = _fixture()
    e = ens.CordEnsemble(z, p)
    print("cord ensemble: all 384 runs", flush=True)
    got = e.run()
    assert len(got) == 384
    bad = [(mc, ens.mean_hz(x), want[mc]) for mc, x in got.items() if ens.mean_hz(x) != want[mc]]
    assert not bad, bad[:10]
    e.release()


def test_lanes_match_checker():
    """The sample as lanes of one ensemble per commissural group."""
    z, p, want = _fixture()
    e = ens.CordEnsemble(z, p)
    print(f"cord ensemble as lanes: {len(SAMPLE)} runs", flush=True)
    got = e.run_lanes(SAMPLE)
    for m, c in SAMPLE:
        ref = orc.cord_lane(z, p, m, c, 1000)
        assert np.array_equal(got[(m, c)], ref), (m, c)
    e.release()


def test_all_runs_as_lanes_reproduce_published_readings():
    z, p, want = _fixture()
    e = ens.CordEnsemble(z, p)
    print("cord ensemble as lanes: all 384 runs", flush=True)
    got = e.run_lanes()
    assert len(got) == 384
    bad = [(mc, ens.mean_hz(x), want[mc]) for mc, x in got.items() if ens.mean_hz(x) != want[mc]]
    assert not bad, bad[:10]
    e.release()


def test_loop_set_twice_runs_like_once():
    """Setting the pool loop again frees the old loop and its captured step
    graphs; the run then equals the checker's."""
    z, p, want = _fixture()
    e = ens.CordEnsemble(z, p)
    m, c = 8, 0
    b = e.start(m, c, steps=200)
    b.pool_counts()
    cfg = ens.config(m, c)
    b.reset()
    b.set_rng(cfg["seed"])
    b.set_pool_feedback(e.pool, e.proprio, 20.0, float(np.float32(0.6)), cfg["pdelay"], 1000, n_pools=6)
    drive = np.zeros(e.N, np.float32)
    drive[e.role == 2] = cfg["cmd"]
    drive[e.is_prop] = np.float32(20.0)
    b.set_array("drive", drive)
    print("cord ensemble: loop set twice", flush=True)
    b.run_async(1000)
    got = b.pool_counts()
    assert np.array_equal(got, orc.cord_lane(z, p, m, c, 1000))
    b.destroy()
    e.release()
