"""The optic lobe and the CNS as one model on the GPU (VisualChainHip, over
fb_fv_* and fb_diff_*) against the CPU checker's chain (src/visdiff.js):

- the shipped coupling (deadband 0.02, cap 200) with a hard threshold and
  Poisson drive: the per-epoch rates, the recorded deviations and the CNS
  spike tape bit-identical; the join's adjoint (dLdV of both eyes and the
  coupling gain's gradient) bit-identical when fed the same drive gradient;
  the optic-lobe and luminance gradients of the whole chain to the CNS
  adjoint's own agreement;
- the soft coupling with the drive's expectation and a soft threshold: rates
  (with the zero-drive members of the driven set) and gradients."""
import importlib

import numpy as np
import pytest

import oracle as orc
from test_visual import Case, STEPS, lum_fn

pytestmark = pytest.mark.gpu
visual = importlib.import_module("fly-brain_amd.visual")


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(1e-30, np.abs(b).max()))


def _pair(case, params, coupling, drive_soft):
    o = orc.VisualChainOracle(case.cns, case.fv, case.map, {**case.pr.p, **params}, coupling=coupling, fvDt=5.0,
                              driveSoft=drive_soft)
    g = visual.VisualChainHip(case.cns, case.fv, case.map,
                              {**case.pr.p, **params, **coupling, "fvDt": 5.0, "driveSoft": drive_soft})
    for ch in (o, g):
        case.pr.apply(ch.net)
        ch.vRest = case.vrest.copy()
        ch.eyeV0 = [v.copy() for v in case.v0]
    return o, g


def test_shipped_coupling_hard_chain_matches_oracle():
    coupling = {"coupling": "hard", "gain": 250.0, "dead": 0.02, "cap": 200.0}
    case = Case(coupling, rest_shift=0.05)
    params = {"soft": False, "tRef": 2.0, "checkpoint": 7}
    print("visual chain: shipped coupling, hard threshold, Poisson drive", flush=True)
    o, g = _pair(case, params, coupling, False)
    Lo = o.forward(STEPS, seed=2, lum=lum_fn(24))
    Lg = g.forward(STEPS, seed=2, lum=lum_fn(24))
    assert np.array_equal(g.sched.view(np.uint32), o.sched.view(np.uint32))
    assert np.array_equal(g.aAt.view(np.uint32), o.aAt.view(np.uint32))
    st = g.stats()
    assert 0 < st["live"] < 1 and st["belowDeadband"] > 0
    so, sg = o.net.spikes(), g.net.spikes()
    for a, b in zip(so, sg):
        assert np.array_equal(a, b)
    assert so[1].size > 50
    assert Lg == pytest.approx(Lo, rel=1e-12)
    go = o.backward(case.pr.dL, lum=True)
    gg = g.backward(case.pr.dL, lum=True)
    assert rel_err(gg["drive"], go["drive"]) < 1e-5
    # the join alone, on the same drive gradient: bit-identical
    dLo, gaino, liveo = o.coupling_backward(go["drive"])
    dLg, gaing, liveg = g.coupling_backward(go["drive"])
    for s in range(2):
        assert np.array_equal(dLg[s].view(np.uint32), dLo[s].view(np.uint32)), s
    assert gaing == gaino and liveg == liveo
    assert abs(gaino) > 0
    # the whole chain: as close as the CNS adjoints
    assert gg["gain"] == pytest.approx(go["gain"], rel=1e-4)
    for k in ("weight", "bias", "kdt"):
        assert rel_err(gg["fv"][k], go["fv"][k]) < 1e-4, k
    for s in range(2):
        assert rel_err(gg["lum"][s], go["lum"][s]) < 1e-4, s
    assert np.count_nonzero(gg["fv"]["weight"]) > 0
    g.destroy()


def test_soft_coupling_chain_matches_oracle():
    coupling = {"coupling": "soft", "gain": 250.0, "dead": 0.02, "cap": 200.0}
    case = Case(coupling)
    print("visual chain: soft coupling, drive as expectation", flush=True)
    o, g = _pair(case, {}, coupling, True)
    Lo = o.forward(STEPS, seed=1, lum=lum_fn(24))
    Lg = g.forward(STEPS, seed=1, lum=lum_fn(24))
    # rates: binary64 exp and sqrt through numpy against the C library's,
    # so equal to a binary32 ulp; the zero-drive members agree exactly
    assert np.array_equal(np.signbit(g.sched) & (g.sched == 0), np.signbit(o.sched) & (o.sched == 0))
    assert np.count_nonzero(np.signbit(o.sched) & (o.sched == 0)) > 0
    assert np.allclose(g.sched, o.sched, rtol=2e-7, atol=0)
    assert Lg == pytest.approx(Lo, rel=1e-6)
    go = o.backward(case.pr.dL, lum=True)
    gg = g.backward(case.pr.dL, lum=True)
    for k in ("wSyn", "inhGain", "vThresh"):
        assert gg["params"][k] == pytest.approx(go["params"][k], rel=1e-4), k
    assert gg["gain"] == pytest.approx(go["gain"], rel=1e-4)
    for k in ("weight", "bias", "kdt"):
        assert rel_err(gg["fv"][k], go["fv"][k]) < 1e-4, k
    assert gg["live"] == go["live"]
    g.destroy()
