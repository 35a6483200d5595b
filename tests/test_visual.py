"""The optic lobe and the CNS as one differentiable model (src/visdiff.js) on
the CPU checker: the chain's adjoint against central differences of its own
smooth forward -- soft threshold, the drive as its expectation (driveSoft),
and either the soft coupling or the hard one without its deadband -- the
check scripts/vis_grad_check.mjs runs on the reference.  Rows: the CNS
scalars and the coupling gain through the join, optic-lobe biases and a
weight (from a fixed start state), and luminance pixels.  Plus the
bookkeeping: the hard coupling's live band and stats(), and a positive rate
too small for binary32 keeping its neuron in the driven set."""
import numpy as np
import pytest

import diffcases as dc
import oracle as orc
from test_flyvis import small_model

STEPS = 120


def lum_fn(ncol, bump=None):
    """a bright bar moving across the columns (scripts/vis_grad_check.mjs:126-127),
    with bump = (eye, epoch, col, delta) added to one pixel"""
    def lum(eye, t):
        c = np.arange(ncol)
        a = (0.5 + 0.45 * np.sin(2 * np.pi * (c / ncol - (1.7 if eye else 1.3) * t))).astype(np.float32)
        if bump is not None and eye == bump[0] and round(t / 0.02) == bump[1]:
            a[bump[2]] += np.float32(bump[3])
        return a
    return lum


class Case:
    """``rest_shift``: vRest is pinned this far below the settled activity
    (vRest is an input the adjoint does not differentiate, src/visdiff.js:
    27-31), so that read nodes the stimulus barely moves sit inside the
    coupling's live band instead of on its kink at rest."""

    def __init__(self, coupling, npair=60, seed=3, rest_shift=0.0):
        self.pr = dc.Problem(N=300, seed=seed, drive_every=0, over={"wSyn": 1.2})
        self.fv = small_model(seed=seed, N=500, nCol=24)
        self.fv["bias"] = np.abs(self.fv["bias"]).astype(np.float32)          # an active lobe
        N = self.pr.g.N
        half = npair // 2
        node0 = self.fv["N"] - npair
        self.map = {"eyes": {"L": {"pairs": [[k, node0 + k] for k in range(half)]},
                             "R": {"pairs": [[k, node0 + k] for k in range(half, npair)]}}}
        self.cns = {"N": N, "indptr": self.pr.g.indptr, "indices": self.pr.g.indices,
                    "counts": self.pr.g.counts, "sign": self.pr.sign}
        self.coupling = coupling
        self.nodes = np.arange(node0, node0 + npair)
        ch = self.chain()
        self.vrest = (ch.settle() - np.float32(rest_shift)).astype(np.float32)
        self.v0 = [v.copy() for v in ch.eyeV0]

    def chain(self, params=None, coupling=None, fv=None):
        ch = orc.VisualChainOracle(self.cns, fv or self.fv, self.map, {**self.pr.p, **(params or {})},
                                   coupling={**self.coupling, **(coupling or {})}, fvDt=5.0, driveSoft=True)
        self.pr.apply(ch.net)
        if getattr(self, "vrest", None) is not None:
            ch.vRest = self.vrest
            ch.eyeV0 = [v.copy() for v in self.v0]
        return ch

    def loss(self, params=None, coupling=None, fv=None, bump=None):
        ch = self.chain(params, coupling, fv)
        return ch.forward(STEPS, seed=1, lum=lum_fn(24, bump), record=False)


def _rows(case, lobe_rows):
    ch = case.chain()
    L = ch.forward(STEPS, seed=1, lum=lum_fn(24))
    g = ch.backward(case.pr.dL, lum=True)
    rows = []
    for k in ("wSyn", "inhGain", "vThresh", "adaptInc"):
        h = dc.H[k]
        p0 = case.pr.p[k]
        fd = (case.loss({k: p0 + h}) - case.loss({k: p0 - h})) / (2 * h)
        rows.append((k, g["params"][k], fd, h))
    h = 0.5
    gain = case.coupling["gain"]
    rows.append(("gain", g["gain"], (case.loss(coupling={"gain": gain + h})
                                     - case.loss(coupling={"gain": gain - h})) / (2 * h), h))
    if not lobe_rows:
        return L, g, rows
    # optic-lobe biases of the read nodes with the largest gradients, and the
    # weight with the largest gradient (the start state is held fixed)
    for i in case.nodes[np.argsort(-np.abs(g["fv"]["bias"][case.nodes]))[:3]]:
        h = 2e-3
        up, dn = case.fv["bias"].copy(), case.fv["bias"].copy()
        up[i] += h
        dn[i] -= h
        rows.append((f"fv bias[{i}]", g["fv"]["bias"][i],
                     (case.loss(fv={**case.fv, "bias": up}) - case.loss(fv={**case.fv, "bias": dn})) / (2 * h), h))
    k = int(np.argmax(np.abs(g["fv"]["weight"])))
    h = 2e-3
    up, dn = case.fv["weight"].copy(), case.fv["weight"].copy()
    up[k] += h
    dn[k] -= h
    rows.append((f"fv weight[{k}]", g["fv"]["weight"][k],
                 (case.loss(fv={**case.fv, "weight": up}) - case.loss(fv={**case.fv, "weight": dn})) / (2 * h), h))
    for eye in (0, 1):
        t, c = np.unravel_index(int(np.argmax(np.abs(g["lum"][eye]))), g["lum"][eye].shape)
        h = 2e-2
        rows.append((f"lum[{eye},{t},{c}]", g["lum"][eye][t, c],
                     (case.loss(bump=(eye, t, c, h)) - case.loss(bump=(eye, t, c, -h))) / (2 * h), h))
    return L, g, rows


def _judge(L, rows, min_rows):
    """scripts/vis_grad_check.mjs:146-156: tolerance 2e-3 |fd| + the f32 loss
    floor 4e-6 |L| / h; rows under three floors are no evidence either way."""
    checked = 0
    for name, a, fd, h in rows:
        floor = 4e-6 * abs(L) / h
        if abs(fd) < 3 * floor:
            continue
        checked += 1
        assert abs(a - fd) <= 2e-3 * abs(fd) + floor, (name, a, fd)
    assert checked >= min_rows, [(r[0], r[2]) for r in rows]


@pytest.mark.parametrize("coupling,lobe_rows,min_rows,shift", [
    ({"coupling": "soft", "gain": 250.0, "dead": 0.02, "cap": 200.0}, False, 4, 0.0),
    ({"coupling": "hard", "gain": 250.0, "dead": 0.0, "cap": 1e9}, True, 9, 0.1),
], ids=["soft", "hard-no-deadband"])
def test_chain_adjoint_matches_finite_differences(oracle_built, coupling, lobe_rows, min_rows, shift):
    """The soft coupling's gate is a logistic of slope 8000 per unit of
    deviation: a step in an optic-lobe parameter or a pixel large enough to
    clear the loss floor carries a pair across the whole gate, so through it
    only the CNS scalars and the gain are differenced (as
    scripts/vis_grad_check.mjs does); the hard coupling without its deadband
    is smooth between the rectifiers' kinks and takes the optic-lobe rows."""
    case = Case(coupling, rest_shift=shift)
    L, g, rows = _rows(case, lobe_rows)
    assert g["live"] > 0.05
    _judge(L, rows, min_rows)


def test_hard_coupling_bookkeeping(oracle_built):
    """The shipped coupling: rate 0 at or below the deadband, gain * a up to
    the cap; its derivative gain inside the band and 0 outside; the live
    fraction and stats() count the same band."""
    case = Case({"coupling": "hard", "gain": 250.0, "dead": 0.02, "cap": 200.0})
    ch = case.chain()
    ch.forward(STEPS, seed=1, lum=lum_fn(24))
    for a in (-0.1, 0.0, 0.02, 0.0200001, 0.5, 0.8, 0.79):
        r = ch._rate(a)
        assert r == (0.0 if a <= 0.02 else min(200.0, 250.0 * a))
        assert ch._drate(a) == (250.0 if 0.02 < a and 250.0 * a < 200.0 else 0.0)
    a = ch.aAt.astype(np.float64).reshape(-1)
    band = (a > 0.02) & (250.0 * a < 200.0)
    gD = np.ones((ch.epochs, ch.nD), np.float32)
    _, gain, live = ch.coupling_backward(gD)
    assert live == pytest.approx(band.mean())
    assert gain == pytest.approx(a[band].sum(), rel=1e-12)
    assert band.any() and (~band).any()


def test_tiny_positive_rate_keeps_the_neuron_driven(oracle_built):
    """setDriveOne tests the binary64 rate (src/lifdiff.js:207): a soft-coupled
    pair at rest has a positive rate far below binary32's range, so its neuron
    is in the driven set with a zero drive and takes a Poisson draw every
    step.  Scheduled as -0.0 the run equals one whose driven set holds that
    neuron at drive 0; scheduled as +0.0, one whose set leaves it out."""
    pr = dc.Problem(N=200, seed=4, drive_every=0, over={"soft": False})
    idx = np.array([5, 9, 17], np.int32)

    def scheduled(val):
        net = pr.oracle()
        net.set_drive_schedule(idx, 50, np.array([[150.0, val, 150.0]], np.float32), drive_soft=False)
        net.forward(50, seed=3, record=True)
        return net.rng, net.spikes()[1].tolist()

    def static(members):
        net = pr.oracle()
        drive = np.zeros(pr.g.N, np.float32)
        drive[[5, 17]] = 150.0
        net.set_input("drive", drive)
        net.set_driven(np.array(members, np.int32))
        net.forward(50, seed=3, record=True)
        return net.rng, net.spikes()[1].tolist()

    assert scheduled(-0.0) == static([5, 9, 17])
    assert scheduled(0.0) == static([5, 17])
    assert scheduled(-0.0)[0] != scheduled(0.0)[0]


@pytest.mark.parametrize("coupling", [
    {"coupling": "hard", "gain": 250.0, "dead": 0.02, "cap": 200.0},
    {"coupling": "soft", "gain": 250.0, "dead": 0.02, "cap": 200.0},
    {"coupling": "hard", "gain": 180.0, "dead": 0.0, "cap": 1e9},
], ids=["hard", "soft", "hard-no-deadband"])
def test_product_coupling_matches_restatement(coupling):
    """The product's vectorised coupling (fly-brain_amd/visual.py) against the
    checker's pair-by-pair restatement: the binary32 drive each rate stores
    (with -0.0 for a positive rate binary32 cannot hold, +0.0 for a zero or
    negative product as Math.max(0, -0) gives), the slope and the gain slope,
    across the deadband, the gate's overflow, the cap and rest."""
    import importlib
    vis = importlib.import_module("fly-brain_amd.visual")
    o = orc.VisualChainOracle.__new__(orc.VisualChainOracle)
    o.c = {"coupling": "hard", **coupling}
    edge = [coupling["dead"] + d for d in (-1e-3, -1e-4, -1e-9, 0.0, 1e-9, 1e-4, 1e-3)]
    a = np.concatenate([np.linspace(-0.3, 1.2, 30001), edge, [coupling["cap"] / coupling["gain"]], [0.0, -0.0]])
    r = vis._drive_value(vis.coupling_rate(a, coupling))
    ro = []
    for x in a:
        v = o._rate(float(x))
        v32 = np.float32(v)
        ro.append(np.float32(-0.0) if (v > 0 and v32 == 0) else v32)
    assert np.array_equal(r.view(np.uint32), np.array(ro, np.float32).view(np.uint32))
    d = vis.coupling_drate(a, coupling)
    dg = vis.coupling_drate_gain(a, coupling)
    do = np.array([o._drate(float(x)) for x in a])
    dgo = np.array([o._drate_gain(float(x)) for x in a])
    assert np.allclose(d, do, rtol=1e-14, atol=0) and np.allclose(dg, dgo, rtol=1e-14, atol=0)
    if coupling["coupling"] == "soft":
        assert np.count_nonzero(np.signbit(r) & (r == 0)) > 0       # some rates are too small for binary32
