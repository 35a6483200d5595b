"""The flyvis optic-lobe step (src/wasm/lif.c:82-93) on the CPU checker: the
restatement (oracle fbo_fv_step) is bit-identical to the reference's own
fv_step compiled natively (oracle/_ref), on random models and on the real
model where the reference's data exist; and it follows the closed form of a
node without inputs from other nodes."""
import importlib
import os

import numpy as np
import pytest

import oracle as orc

VISION = "/root/reference/public/vision"
flyvis = importlib.import_module("fly-brain_amd.flyvis")


def random_model(N=2000, deg=30, seed=1):
    rng = np.random.default_rng(seed)
    d = rng.poisson(deg, N)
    indptr = np.concatenate([[0], np.cumsum(d)]).astype(np.int32)
    target = rng.integers(0, N, int(indptr[-1])).astype(np.int32)
    weight = rng.normal(0, 0.08, int(indptr[-1])).astype(np.float32)
    bias = rng.normal(0.1, 0.3, N).astype(np.float32)
    tau = rng.uniform(0.005, 0.2, N).astype(np.float32)
    return {"N": N, "E": int(indptr[-1]), "indptr": indptr, "target": target, "weight": weight, "bias": bias,
            "tau": tau, "dt": 0.02}


def run_checker(kind, m, xs, steps_per_x=5):
    v = m["bias"].astype(np.float32).copy()
    acc = np.zeros(m["N"], np.float32)
    kdt = orc.fv_kdt(m["tau"], m["dt"])
    for x in xs:
        for _ in range(steps_per_x):
            orc.fv_step(kind, m["bias"], kdt, m["indptr"], m["target"], m["weight"], v, acc, x)
    return v, acc


def _inputs(N, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0, 1, N).astype(np.float32) for _ in range(n)]


@pytest.mark.skipif(not orc.available("ref"), reason="oracle/_ref is not built")
def test_restatement_matches_reference_fv_step_on_random_models(oracle_built):
    for seed in (1, 2):
        m = random_model(seed=seed)
        xs = _inputs(m["N"], 6, seed)
        vo, ao = run_checker("oracle", m, xs)
        vr, ar = run_checker("ref", m, xs)
        assert np.array_equal(vo.view(np.uint32), vr.view(np.uint32))
        assert np.array_equal(ao.view(np.uint32), ar.view(np.uint32))


@pytest.mark.skipif(not (orc.available("ref") and os.path.exists(os.path.join(VISION, "flyvis.bin"))),
                    reason="oracle/_ref or the reference's vision data are absent")
def test_restatement_matches_reference_fv_step_on_the_real_model(oracle_built):
    m = flyvis.parse_flyvis(VISION)
    assert m["N"] == 45669 and m["E"] == 1513231
    rng = np.random.default_rng(3)
    xs = []
    for _ in range(4):
        x = np.zeros(m["N"], np.float32)
        lum = rng.uniform(0, 1, 721).astype(np.float32)
        for ix in m["inputIdx"].values():
            x[ix] = lum[: ix.size]
        xs.append(x)
    vo, ao = run_checker("oracle", m, xs, steps_per_x=10)
    vr, ar = run_checker("ref", m, xs, steps_per_x=10)
    assert np.array_equal(vo.view(np.uint32), vr.view(np.uint32))
    assert np.array_equal(ao.view(np.uint32), ar.view(np.uint32))
    assert np.isfinite(vo).all() and (vo != m["bias"]).any()


def test_lone_node_closed_form(oracle_built):
    """No edges: v_{n+1} = v_n + kdt (-v_n + bias + x), so v_n -> bias + x
    geometrically with ratio 1 - kdt."""
    N = 3
    m = {"N": N, "E": 0, "indptr": np.zeros(N + 1, np.int32), "target": np.zeros(0, np.int32),
         "weight": np.zeros(0, np.float32), "bias": np.array([0.2, -0.1, 0.0], np.float32),
         "tau": np.array([0.1, 0.04, 0.01], np.float32), "dt": 0.02}
    x = np.array([0.5, 0.3, 0.9], np.float32)
    v, acc = run_checker("oracle", m, [x], steps_per_x=40)
    kdt = orc.fv_kdt(m["tau"]).astype(np.float64)
    target = m["bias"].astype(np.float64) + x
    want = target + (m["bias"] - target) * (1 - kdt) ** 40
    assert np.allclose(v, want, rtol=0, atol=2e-6)
    assert kdt[2] == 1.0 and v[2] == np.float32(0.9)   # tau below dt: kdt = 1, v = bias + x at once
    assert (acc == 0).all()


def test_rectified_sources_only(oracle_built):
    """A source with v <= 0 adds nothing; one with v > 0 adds w * v."""
    m = {"N": 3, "E": 2, "indptr": np.array([0, 1, 2, 2], np.int32), "target": np.array([2, 2], np.int32),
         "weight": np.array([2.0, 5.0], np.float32), "bias": np.array([0.5, -0.5, 0.0], np.float32),
         "tau": np.full(3, 1.0, np.float32), "dt": 0.02}
    v, acc = run_checker("oracle", m, [np.zeros(3, np.float32)], steps_per_x=1)
    assert acc[2] == np.float32(2.0) * np.float32(0.5)
