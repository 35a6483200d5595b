"""The flyvis optic-lobe step on the GPU (fb_fv_*, FlyVisHip) against the CPU
restatement of fv_step (src/wasm/lif.c:82-93): activity and synaptic input
bit-identical after every block of steps, two eyes sharing one model, on
random models and on the reference's model (from tests/golden/_vision/,
copied by tests/golden/make_vision_fixture.py)."""
import importlib
import os

import numpy as np
import pytest

import oracle as orc
from test_flyvis import random_model

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
VISION = os.path.join(HERE, "golden", "_vision")
flyvis = importlib.import_module("fly-brain_amd.flyvis")


def _check(m, frames, steps_per_frame, use_lum):
    eye0 = flyvis.FlyVisHip(m)
    eye1 = flyvis.FlyVisHip(shared=eye0.model)
    kdt = orc.fv_kdt(m["tau"], m["dt"])
    v = [m["bias"].astype(np.float32).copy() for _ in range(2)]
    acc = [np.zeros(m["N"], np.float32) for _ in range(2)]
    for f, frame in enumerate(frames):
        for e, eye in enumerate((eye0, eye1)):
            inp = frame if e == 0 else frame[::-1].copy()
            if use_lum:
                eye.setInput(inp)
                x = eye.x.copy()
            else:
                eye.set_x(inp)
                x = inp
            eye.step(steps_per_frame)
            for _ in range(steps_per_frame):
                orc.fv_step("oracle", m["bias"], kdt, m["indptr"], m["target"], m["weight"], v[e], acc[e], x)
            assert np.array_equal(eye.v.view(np.uint32), v[e].view(np.uint32)), (f, e)
            assert np.array_equal(eye.acc.view(np.uint32), acc[e].view(np.uint32)), (f, e)
    eye0.reset()
    assert np.array_equal(eye0.v, m["bias"])
    eye0.destroy()
    eye1.destroy()


@pytest.mark.parametrize("seed", [1, 2])
def test_random_model_matches_oracle(seed):
    m = random_model(N=3000, deg=40, seed=seed)
    rng = np.random.default_rng(seed)
    print(f"flyvis random model seed={seed}", flush=True)
    _check(m, [rng.uniform(0, 1, m["N"]).astype(np.float32) for _ in range(5)], 7, use_lum=False)


def test_reference_model_matches_oracle():
    if not os.path.exists(os.path.join(VISION, "flyvis.bin")):
        pytest.skip("the vision fixture was not built (needs the reference's data files)")
    m = flyvis.parse_flyvis(VISION)
    rng = np.random.default_rng(7)
    print("flyvis reference model", flush=True)
    _check(m, [rng.uniform(0, 1, 721).astype(np.float32) for _ in range(4)], 10, use_lum=True)
