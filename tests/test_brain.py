"""The whole calibrated brain (fly-brain_amd/connectome.py, 165,122 neurons)
against the reference's own published answer: the sugar trace of its twin
audit, four seeds of the shipped kernel (public/data/twin_audit.json, written
by scripts/twin_audit.mjs:159-172), replayed on the lif.c oracle.  The total
pins every step of the build: connectome files, input scale, sensory mask with
the eye-coupled neurons, octopaminergic sign removal, writeGraph's bake, class
physiology, the fed octopamine tone, LIFWasm's seeding and drive order.
Skipped where the reference's data are absent."""
import importlib
import json
import os

import numpy as np
import pytest

from oracle import OracleBrain

DATA = "/root/reference/public/data"
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(DATA, "graph_w3.bin")),
                                reason="the reference's data files are absent")
C = importlib.import_module("fly-brain_amd.connectome")

# public/data/twin_audit.json, traces[0] ("sugar", 4 seeds, 0.7 s): the mean
# over seeds 1-4 of the brain's total spike count, and the mean spikes per
# neuron of three reported types
SUGAR_TOTAL_MEAN = 448344.0
SUGAR_TYPE_MEANS = {"MN9": 20.375, "GNG232": 7.25, "DNge080": 9.375}
SEEDS = (1, 2, 3, 4)


class _G:
    pass


def sugar_counts(b, assay, seed):
    """One seed of the sugar trace on the oracle: spike counts after 1,400 steps."""
    g = _G()
    g.N, g.indptr, g.indices = b["N"], b["indptr"], b["indices"]
    o = OracleBrain(g, b["params"], seed=seed, W=b["W"], sign=b["sign"])
    for i in np.nonzero(b["thr"])[0]:
        o.setThr(int(i), float(b["thr"][i]))
    lam = np.nonzero(b["bias"])[0]
    if lam.size:
        o.setBias(lam, float(b["bias"][lam[0]]))
    o.setDrive(assay["orn"], assay["orn_hz"])
    for i in assay["sugar"]:
        o.setDriveOne(int(i), assay["sugar_hz"])
    o.run(assay["on_steps"], raster_cap=1)
    for i in assay["sugar"]:
        o.setDriveOne(int(i), 0.0)
    o.run(assay["steps"] - assay["on_steps"], raster_cap=1)
    return o.spikeCount.copy()


def trace_summary(b, conn):
    assay = C.sugar_assay(conn)
    counts = [sugar_counts(b, assay, s) for s in SEEDS]
    total = sum(int(c.sum()) for c in counts) / len(SEEDS)
    means = {}
    for t in SUGAR_TYPE_MEANS:
        ix = C.by_type(conn, t)
        means[t] = round(sum(int(c[ix].sum()) for c in counts) / len(SEEDS) / ix.size, 3)
    return total, means


@pytest.fixture(scope="module")
def brain():
    return C.calibrated_brain(DATA)


def test_calibrated_brain_reproduces_the_sugar_trace(brain):
    total, means = trace_summary(brain, brain["conn"])
    assert total == SUGAR_TOTAL_MEAN
    assert means == SUGAR_TYPE_MEANS


def test_build_details(brain):
    """Sizes and parameters of the build (public/data/meta.json, brain_params.json)."""
    assert brain["N"] == 165122 and brain["W"].size == 10511038
    p = brain["params"]
    assert p["dt"] == 0.5 and p["coba"] is True and p["neuromod"] is True and p["minSyn"] == 6
    conn = brain["conn"]
    # octopaminergic sources carry no fast synapses
    oa = np.array([str(t).startswith("OA-") for t in conn["meta"]["types"]])
    assert np.all(brain["sign"][oa] == 0) and oa.sum() > 0
    # Kenyon cells start at kcThreshold unless an octopamine target shifted them lower
    kc = np.array([conn["meta"]["classes"][c] == "Kenyon_Cell" for c in conn["cls"]])
    assert np.all(brain["thr"][kc] <= np.float32(p["kcThreshold"]))
    assert np.all(brain["thr"][kc] > np.float32(p["kcThreshold"]) - 2)


def test_a_planted_tone_error_misses_the_total(brain):
    """Without the octopamine tone's threshold shift the total moves."""
    conn = brain["conn"]
    thr, _ = C.class_physiology(conn, brain["params"])
    b = dict(brain, thr=thr)
    total, _ = trace_summary(b, conn)
    assert total != SUGAR_TOTAL_MEAN


def test_a_planted_mask_error_misses_the_total(brain):
    """Without the eye-coupled neurons in the sensory mask the total moves."""
    conn, p = brain["conn"], brain["params"]
    in_scale, sensory, _ = C.brain_scales(conn, conn["size"], p)
    W, sign = C.write_graph(conn["N"], conn["indptr"], conn["indices"], conn["weights"],
                            C.modulatory_sign(conn, conn["ntsign"], p), p, in_scale, sensory)
    total, _ = trace_summary(dict(brain, W=W, sign=sign), conn)
    assert total != SUGAR_TOTAL_MEAN


def test_compact_form_round_trips(tmp_path, brain):
    path = str(tmp_path / "brain.npz")
    C.save_compact(path, DATA)
    z = C.load_compact(path)
    for k in ("indptr", "indices"):
        assert np.array_equal(z[k], brain[k])
    for k in ("W", "sign", "thr", "bias"):
        assert np.array_equal(z[k].view(np.uint32), brain[k].view(np.uint32))


def test_type_gains():
    """src/brainmodel.js:63-100 on a toy connectome: exact names, regular
    expressions over the whole name, class gains, a gain table; None without
    any (the calibrated parameters have none)."""
    conn = {"N": 5, "meta": {"types": ["T4a", "T4b", "MN9", None, "T4ab"], "classes": ["x", "y"]},
            "cls": np.array([0, 0, 1, 1, 0])}
    assert C.type_gains(conn, {}) is None
    assert C.type_gains(conn, {"typeGain": {"MN9": 1}}) is None
    g = C.type_gains(conn, {"typeGain": {"T4a|T4b": 2.0, "MN9": 0.5}, "classGain": {"y": 3.0},
                            "neuronGainTable": {"4": 0.0, "0": np.log(2.0)}})
    want = np.array([2.0 * 2.0, 2.0, 0.5 * 3.0, 3.0, 1.0], dtype=np.float32)
    assert np.allclose(g, want, rtol=1e-6) and g.dtype == np.float32
    # a lognormal individual: mean log gain near 0, spread near sigma
    big = {"N": 20000, "meta": {"types": ["a"] * 20000, "classes": ["x"]}, "cls": np.zeros(20000, int)}
    lg = np.log(C.type_gains(big, {"neuronGain": {"sigma": 0.2, "seed": 3}}).astype(np.float64))
    assert abs(lg.mean()) < 0.01 and abs(lg.std() - 0.2) < 0.01
