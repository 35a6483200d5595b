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


# ---------------------------------------------------------------- the adjoint (src/flyvisdiff.js)

def small_model(seed=5, N=300, nCol=24):
    m = random_model(N=N, deg=10, seed=seed)
    m["weight"] = (m["weight"] * 2.5).astype(np.float32)
    rng = np.random.default_rng(seed)
    nodes = rng.permutation(N)[: 2 * nCol]
    m["inputIdx"] = {"R1": nodes[:nCol], "R2": nodes[nCol:]}
    return m


def _loss64(m, inNode, inCol, lums, dLdV, bias=None, kdt=None, weight=None):
    """L = sum_t dLdV_t . v_t of the rate network run in f64 from v = bias of
    the model (the start state is an input, held fixed while a parameter
    moves)."""
    N = int(m["N"])
    bias = m["bias"].astype(np.float64) if bias is None else bias
    w = m["weight"].astype(np.float64) if weight is None else weight
    src = np.repeat(np.arange(N), np.diff(m["indptr"]))
    v = m["bias"].astype(np.float64).copy()
    L, vs = 0.0, []
    for t, lum in enumerate(lums):
        x = np.zeros(N)
        x[inNode] = np.asarray(lum, np.float64)[inCol]
        acc = np.zeros(N)
        np.add.at(acc, m["target"], w * np.maximum(v, 0.0)[src])
        v = v + kdt * (-v + bias + acc + x)
        vs.append(v)
        L += float(dLdV[t].astype(np.float64) @ v)
    return L, vs


def test_flyvis_adjoint_matches_finite_differences(oracle_built):
    """dL/d bias, kdt, weight and luminance of L = sum_t dLdV_t . v_t against
    central differences of the forward (src/flyvisdiff.js:4-6: the network is
    smooth away from the rectifier's kink, so the differences are taken in f64
    with a step far below the distance to any kink; the f64 forward is first
    checked against the binary32 forward the reference runs)."""
    m = small_model()
    rng = np.random.default_rng(6)
    T = 24
    lums = [rng.uniform(0, 1, 24).astype(np.float32) for _ in range(T)]
    dLdV = np.zeros((T, m["N"]), np.float32)
    dLdV[:, rng.choice(m["N"], 40, replace=False)] = rng.normal(0, 1, (T, 40))
    o = orc.FlyVisDiffOracle(m)
    o.start_tape()
    vs32 = []
    for t in range(T):
        o.setInput(lums[t])
        o.step()
        vs32.append(o.v.copy())
    kdt = o.kdt.astype(np.float64)
    L, vs = _loss64(m, o.inNode, o.inCol, lums, dLdV, kdt=kdt)
    assert np.abs(np.array(vs) - np.array(vs32)).max() < 1e-4
    dLum = np.zeros((T, o.nCol), np.float32)
    g = o.backward(dLdV, dLdLum=dLum)
    h = 1e-6

    def fd(**kw):
        """central difference along kw = {name: (value, direction)}"""
        up = {k: (v[0] + h * v[1]) for k, v in kw.items()}
        dn = {k: (v[0] - h * v[1]) for k, v in kw.items()}
        return (_loss64(m, o.inNode, o.inCol, lums, dLdV, kdt=up.pop("kdt", kdt), **up)[0]
                - _loss64(m, o.inNode, o.inCol, lums, dLdV, kdt=dn.pop("kdt", kdt), **dn)[0]) / (2 * h)

    def close(a, f):
        return abs(a - f) <= 1e-4 * abs(f) + 1e-5

    ok = 0
    b64 = m["bias"].astype(np.float64)
    for i in rng.choice(m["N"], 8, replace=False):
        e = np.zeros(m["N"])
        e[i] = 1.0
        f = fd(bias=(b64, e))
        assert close(g["bias"][i], f), ("bias", i, g["bias"][i], f)
        f = fd(kdt=(kdt, e))
        assert close(g["kdt"][i], f), ("kdt", i, g["kdt"][i], f)
        ok += abs(f) > 1e-3
    w64 = m["weight"].astype(np.float64)
    live = np.flatnonzero(g["weight"] != 0)
    for k in rng.choice(live, 8, replace=False):
        e = np.zeros(m["E"])
        e[k] = 1.0
        f = fd(weight=(w64, e))
        assert close(g["weight"][k], f), ("weight", k, g["weight"][k], f)
    for t, c in ((0, 3), (5, 10), (12, 0), (20, 23), (23, 7)):
        lu = [x.astype(np.float64) for x in lums]
        ld = [x.astype(np.float64) for x in lums]
        lu[t][c] += h
        ld[t][c] -= h
        f = (_loss64(m, o.inNode, o.inCol, lu, dLdV, kdt=kdt)[0]
             - _loss64(m, o.inNode, o.inCol, ld, dLdV, kdt=kdt)[0]) / (2 * h)
        assert close(dLum[t, c], f), ("lum", t, c, dLum[t, c], f)
    assert ok >= 3
    assert np.count_nonzero(dLum[:, :o.nCol]) > T   # the input map reaches the loss


def test_flyvis_adjoint_closed_form_lone_node(oracle_built):
    """One node, no edges, loss v_T: dv_T/dbias = 1 - (1 - kdt)^T and
    dv_T/dx_t = kdt (1 - kdt)^(T-1-t)."""
    m = {"N": 1, "E": 0, "indptr": np.zeros(2, np.int32), "target": np.zeros(0, np.int32),
         "weight": np.zeros(0, np.float32), "bias": np.array([0.3], np.float32), "tau": np.array([0.1], np.float32),
         "dt": 0.02, "inputIdx": {"R1": np.array([0])}}
    T = 10
    o = orc.FlyVisDiffOracle(m)
    o.start_tape()
    for t in range(T):
        o.setInput(np.array([0.5], np.float32))
        o.step()
    dLdV = np.zeros((T, 1), np.float32)
    dLdV[-1, 0] = 1.0
    dLum = np.zeros((T, 1), np.float32)
    g = o.backward(dLdV, dLdLum=dLum)
    k = float(o.kdt[0])
    assert g["bias"][0] == pytest.approx(1 - (1 - k) ** T, rel=1e-6)
    for t in range(T):
        assert dLum[t, 0] == pytest.approx(k * (1 - k) ** (T - 1 - t), rel=1e-6)
