"""The differentiable flyvis on the GPU (fb_fv_tape_* / fb_fv_backward,
FlyVisDiffHip) against the CPU restatement of FlyVisDiff
(src/flyvisdiff.js:36-185): the taped forward and the weight, bias, kdt and
luminance gradients are bit-identical, with two eyes on one model adding
into one accumulator as src/visdiff.js:235-238 does; on random models and on
the reference's model (tests/golden/_vision/, copied by
tests/golden/make_vision_fixture.py)."""
import importlib
import os

import numpy as np
import pytest

import oracle as orc
from test_flyvis import small_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
VISION = os.path.join(HERE, "golden", "_vision")
flyvis = importlib.import_module("fly-brain_amd.flyvis")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _case(m, T, n_loss, seed, start=None):
    rng = np.random.default_rng(seed)
    nCol = max(len(a) for a in m["inputIdx"].values())
    lums = [[rng.uniform(0, 1, nCol).astype(np.float32) for _ in range(T)] for _ in range(2)]
    dLdV = []
    for _ in range(2):
        d = np.zeros((T, m["N"]), np.float32)
        d[:, rng.choice(m["N"], n_loss, replace=False)] = rng.normal(0, 1, (T, n_loss))
        dLdV.append(d)
    # the oracle: two eyes, one accumulator
    g_o = {"weight": np.zeros(m["E"], np.float32), "bias": np.zeros(m["N"], np.float32),
           "kdt": np.zeros(m["N"], np.float32)}
    eyes_o, lum_o = [], []
    for e in range(2):
        o = orc.FlyVisDiffOracle(m, kdt=m.get("kdt"))
        if start is not None:
            o.v = start[e].copy()
        o.start_tape()
        for t in range(T):
            o.setInput(lums[e][t])
            o.step()
        eyes_o.append(o)
    for e in range(2):
        lum_o.append(np.zeros((T, eyes_o[e].nCol), np.float32))
        eyes_o[e].backward(dLdV[e], g_o, lum_o[e])
    # the GPU
    eye0 = flyvis.FlyVisDiffHip(m)
    eye1 = flyvis.FlyVisDiffHip(shared=eye0.sharedParts)
    g = {"weight": np.zeros(m["E"], np.float32), "bias": np.zeros(m["N"], np.float32),
         "kdt": np.zeros(m["N"], np.float32)}
    lum_g = []
    for e, eye in enumerate((eye0, eye1)):
        if start is not None:
            eye.set_v(start[e])
        eye.startTape(T)
        for t in range(T):
            eye.setInput(lums[e][t])
            eye.stepTaped()
        assert eye.tape_len == T
        assert np.array_equal(_bits(eye.v), _bits(eyes_o[e].v)), ("forward", e)
    for e, eye in enumerate((eye0, eye1)):
        lum_g.append(np.zeros(T * eye.nCol, np.float32))
        eye.backward(dLdV[e], g, lum_g[e])
    for k in ("weight", "bias", "kdt"):
        assert np.array_equal(_bits(g[k]), _bits(g_o[k])), k
    for e in range(2):
        assert np.array_equal(_bits(lum_g[e]), _bits(lum_o[e].reshape(-1))), ("lum", e)
    assert np.count_nonzero(g["weight"]) > 0 and np.count_nonzero(lum_g[0]) > 0
    return eye0, eye1


@pytest.mark.parametrize("seed", [3, 4])
def test_random_model_gradients_match_oracle(seed):
    m = small_model(seed=seed, N=2500, nCol=60)
    print(f"flyvis adjoint random model seed={seed}", flush=True)
    eye0, eye1 = _case(m, T=30, n_loss=200, seed=seed)
    eye0.destroy()
    eye1.destroy()


def test_hub_nodes_span_several_chunks():
    """A node with 3,000 in-edges and one with 3,000 out-edges: a wave group
    of one node whose edges are staged in several chunks, forward and
    backward."""
    m = small_model(seed=12, N=1500, nCol=30)
    rng = np.random.default_rng(12)
    src = np.repeat(np.arange(m["N"]), np.diff(m["indptr"]))
    src = np.concatenate([src, rng.integers(0, m["N"], 3000), np.full(3000, 11)])
    tgt = np.concatenate([m["target"], np.full(3000, 7), rng.integers(0, m["N"], 3000)])
    w = np.concatenate([m["weight"], rng.normal(0, 0.01, 6000).astype(np.float32)])
    o = np.argsort(src, kind="stable")
    m["target"], m["weight"] = tgt[o].astype(np.int32), w[o].astype(np.float32)
    m["indptr"] = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=m["N"]))]).astype(np.int32)
    m["E"] = int(m["indptr"][-1])
    m["bias"][11] = 0.8                       # the fan-out hub is active
    print("flyvis adjoint: hub nodes", flush=True)
    eye0, eye1 = _case(m, T=10, n_loss=150, seed=12)
    eye0.destroy()
    eye1.destroy()


def test_kdt_given_and_tape_bounds():
    """model.kdt replaces dt / max(tau, dt) (src/flyvisdiff.js:50-53); a start
    state other than rest; taping past the tape's room is refused."""
    m = small_model(seed=8, N=1200, nCol=40)
    m["kdt"] = np.random.default_rng(8).uniform(0.05, 0.9, m["N"]).astype(np.float32)
    start = [np.random.default_rng(9 + e).normal(0.2, 0.5, m["N"]).astype(np.float32) for e in range(2)]
    print("flyvis adjoint: given kdt", flush=True)
    eye0, eye1 = _case(m, T=12, n_loss=100, seed=8, start=start)
    with pytest.raises(RuntimeError, match="tape holds"):
        eye0.stepTaped()
    eye0.startTape(4)                         # a new tape restarts the count
    eye0.stepTaped(4)
    assert eye0.tape_len == 4
    eye0.destroy()
    eye1.destroy()


def test_reference_model_gradients_match_oracle():
    if not os.path.exists(os.path.join(VISION, "flyvis.bin")):
        pytest.skip("the vision fixture was not built (needs the reference's data files)")
    m = flyvis.parse_flyvis(VISION)
    print("flyvis adjoint reference model", flush=True)
    eye0, eye1 = _case(m, T=8, n_loss=600, seed=11)
    eye0.destroy()
    eye1.destroy()
