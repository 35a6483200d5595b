"""Pins the CPU restatement of the differentiable model (oracle/lifdiff_oracle.c).

The reference's own gradient check (scripts/grad_check.mjs) compares the
adjoint with central finite differences of the smooth (soft-threshold) model;
the same comparison, with its steps and tolerance, is made here for the nine
scalar gradients, per-neuron log-gains, per-type gains, single edges and the
filtered-rate readout loss.  Invariants of the per-edge gradient (row and
column sums), replay exactness across checkpoint lengths, and the drive
stream against the lif.c restatement pin what finite differences leave open."""
import importlib

import numpy as np
import pytest

import diffcases as dc

synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

STEPS = 60


@pytest.fixture(scope="module")
def base(oracle_built):
    pr = dc.Problem()
    net = pr.oracle()
    L = net.forward(STEPS, seed=1)
    g = net.backward()
    return pr, L, g


def _loss(pr, **kw):
    return pr.oracle(**kw).forward(STEPS, seed=1, record=False)


@pytest.mark.parametrize("name", P.DIFF_PARAMS)
def test_scalar_gradients_match_finite_differences(base, name):
    pr, L, g = base
    h = dc.H[name]
    fd = (_loss(pr, params={name: pr.p[name] + h}) - _loss(pr, params={name: pr.p[name] - h})) / (2 * h)
    ok, rel = dc.fd_ok(g["params"][name], fd, h, L)
    assert ok, (name, g["params"][name], fd, rel)
    assert g["params"][name] != 0.0


def test_log_gain_gradients_match_finite_differences(base):
    pr, L, g = base
    h = 2e-2
    for i in (0, 1, 2, pr.g.N // 3, pr.g.N // 2):
        up = pr.inputs["logGain"].copy()
        dn = up.copy()
        up[i] += h
        dn[i] -= h
        fd = (_loss(pr, inputs={"logGain": up}) - _loss(pr, inputs={"logGain": dn})) / (2 * h)
        ok, rel = dc.fd_ok(g["logGain"][i], fd, h, L)
        assert ok, (i, g["logGain"][i], fd, rel)


def test_type_gain_gradient_is_the_type_sum_of_log_gain_gradients(base):
    """outScale[i] = gain[type(i)]; d L / d log gain_T = sum over the type of
    gLogGain (src/brainmodel.js:63-100 semantics of a type gain)."""
    pr, L, g = base
    types = np.random.default_rng(5).integers(0, 12, size=pr.g.N)
    h = 1e-3
    for T in (0, 3, 7):
        mine = types == T
        adj = float(g["logGain"][mine].astype(np.float64).sum())
        up = pr.inputs["outScale"].astype(np.float64).copy()
        dn = up.copy()
        up[mine] *= np.exp(h)
        dn[mine] *= np.exp(-h)
        fd = (_loss(pr, inputs={"outScale": up.astype(np.float32)}) -
              _loss(pr, inputs={"outScale": dn.astype(np.float32)})) / (2 * h)
        ok, rel = dc.fd_ok(adj, fd, h, L)
        assert ok, (T, adj, fd, rel)


def test_edge_gradients_match_finite_differences(base):
    pr, L, g = base
    E = pr.g.indices.size
    top = np.argsort(-np.abs(g["edge"]))[:6]
    h = 5e-2
    for j in top:
        up = np.ones(E, np.float32)
        dn = np.ones(E, np.float32)
        up[j] += h
        dn[j] -= h
        fd = (_loss(pr, edge_gain=up) - _loss(pr, edge_gain=dn)) / (2 * h)
        assert abs(g["edge"][j] - fd) <= 2e-3 * abs(fd) + 4e-6 * abs(L) / h, (j, g["edge"][j], fd)


def test_edge_gradient_row_and_column_invariants(base):
    """sum over row(pre) of gEdge = gLogGain[pre]; sum over edges into q of
    gEdge = inScale[q] * gInScale[q] (edge gains 1)."""
    pr, L, g = base
    pre = np.repeat(np.arange(pr.g.N), np.diff(pr.g.indptr.astype(np.int64)))
    row = np.bincount(pre, weights=g["edge"].astype(np.float64), minlength=pr.g.N)
    scale = np.abs(g["logGain"]).max()
    assert np.abs(row - g["logGain"]).max() <= 1e-5 * scale
    col = np.bincount(pr.g.indices, weights=g["edge"].astype(np.float64), minlength=pr.g.N)
    net = pr.oracle()
    net.forward(1, seed=1, record=False)
    ins = net.state("inScale").astype(np.float64)
    want = ins * g["inScale"]
    assert np.abs(col - want).max() <= 1e-5 * np.abs(want).max()


def test_readout_rate_loss_gradients_match_finite_differences(oracle_built):
    """The filtered-rate MSE on readouts, alone (flat spike weights zero)."""
    pr = dc.Problem(seed=13).with_readout(R=40, tau_ms=20.0)
    pr.dL[:] = 0.0
    net = pr.oracle()
    L = net.forward(STEPS, seed=2)
    g = net.backward()
    s, r = net.readout()
    # the loss is the mean squared rate error
    assert L == pytest.approx(float(((r - pr.readout[1][None, :]) ** 2).mean()), rel=1e-12)
    for name in ("wSyn", "vThresh", "adaptInc", "depU", "inhGain"):
        h = dc.H[name]
        fd = (pr.oracle(params={name: pr.p[name] + h}).forward(STEPS, 2, False) -
              pr.oracle(params={name: pr.p[name] - h}).forward(STEPS, 2, False)) / (2 * h)
        ok, rel = dc.fd_ok(g["params"][name], fd, h, L)
        assert ok, (name, g["params"][name], fd, rel)


def test_rate_filter_closed_form(oracle_built):
    """One readout driven by a forced spike train: r_t is the exponential
    filter of s_t / (dt/1000) with a = exp(-dt/tau)."""
    pr = dc.Problem(N=64, deg=0.0, seed=3, over={"soft": False}, drive_every=1, drive_hz=400.0)
    pr.readout = (np.array([5], np.int32), np.array([10.0], np.float32), 20.0)
    net = pr.oracle()
    net.forward(200, seed=4)
    s, r = net.readout()
    a = np.exp(-pr.p["dt"] / 20.0)
    want = np.zeros(200)
    acc = 0.0
    for t in range(200):
        acc = a * acc + (1 - a) * float(s[t, 0]) / (pr.p["dt"] / 1000.0)
        want[t] = acc
    assert s[:, 0].sum() > 3
    assert np.allclose(r[:, 0], want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("soft", [True, False])
def test_gradients_do_not_depend_on_checkpoint_length(oracle_built, soft):
    """Segment replay reproduces the recorded trajectory exactly, so the
    gradients are bit-identical for any checkpoint spacing."""
    out = []
    for K in (16, 7, 60, 1):
        pr = dc.Problem(over={"soft": soft, "checkpoint": K, "tRef": 2.2 if not soft else 0.0})
        net = pr.oracle()
        net.forward(STEPS, seed=1)
        out.append(net.backward())
    for o in out[1:]:
        assert o["params"] == out[0]["params"]
        assert np.array_equal(o["logGain"], out[0]["logGain"])
        assert np.array_equal(o["edge"], out[0]["edge"])


def test_truncation_and_clip_limits(oracle_built):
    pr = dc.Problem()
    net = pr.oracle()
    net.forward(STEPS, seed=1)
    full = net.backward()
    same = net.backward(truncate=STEPS)          # clears only before the last step
    assert same["params"] == full["params"]
    short = net.backward(truncate=5)
    assert short["params"]["vThresh"] != full["params"]["vThresh"]
    pr2 = dc.Problem(over={"adjClip": 1e30})
    n2 = pr2.oracle()
    n2.forward(STEPS, seed=1)
    assert n2.backward()["params"] == full["params"]


def test_hard_forward_drive_stream_matches_lif_c(oracle_built):
    """With no synapses, LIFDiff's hard forward and lif.c draw the same
    xorshift32 stream over the same driven order, so the forced spikes agree
    step by step (src/lifdiff.js:203-213 vs src/wasm/lif.c:49-52)."""
    from oracle import OracleBrain
    N = 200
    g = synth.Graph(N=N, indptr=np.zeros(N + 1, np.uint32), indices=np.zeros(0, np.uint32),
                    counts=np.zeros(0, np.int32), exc=np.ones(N, bool))
    drv = np.random.default_rng(2).permutation(N)[:80].astype(np.int32)
    rates = np.random.default_rng(3).uniform(20, 400, size=80).astype(np.float32)
    pd = P.diff_merged({"dt": 0.1, "tRef": 2.2})
    from oracle import OracleDiff
    od = OracleDiff(N, g.indptr, g.indices, np.zeros(0, np.float32), np.ones(N, np.float32), pd)
    drive = np.zeros(N, np.float32)
    drive[drv] = rates
    od.set_input("drive", drive)
    od.set_driven(drv)
    od.set_loss(np.zeros(N, np.float32))
    od.forward(3000, seed=7)
    off, idx, amp, forced = od.spikes()
    pl = P.config_params(tRef=2.2)
    ob = OracleBrain(g, pl, seed=7, W=np.zeros(0, np.float32), sign=np.ones(N, np.float32))
    for i, r in zip(drv, rates):
        ob.setDriveOne(int(i), float(r))
    nf, ras, _ = ob.run(3000, raster_cap=3000 * N)
    assert forced.all() and np.array_equal(np.diff(off), nf)
    assert np.array_equal(idx, ras)
    assert od.rng == ob.rng


# ---------------------------------------------------------------- closed forms of the forward pass

def _lone(N, edges=(), counts=(), sign=None, over=None):
    from oracle import OracleDiff
    indptr = np.zeros(N + 1, np.uint32)
    src = np.array([e[0] for e in edges], np.int64)
    for s_ in src:
        indptr[s_ + 1:] += 1
    idx = np.array([e[1] for e in edges], np.uint32)
    p = P.diff_merged({"dt": 0.1, **(over or {})})
    net = OracleDiff(N, indptr, idx, np.array(counts, np.float32),
                     np.ones(N, np.float32) if sign is None else np.array(sign, np.float32), p)
    net.set_loss(np.zeros(N, np.float32))
    return net, p


def test_membrane_closed_form_under_constant_bias():
    """No input, laminaBias b on the neuron: v_n = vRest + b (1 - (1 - kM)^n)
    (src/lifdiff.js:320-321), below threshold."""
    net, p = _lone(1, over={"laminaBias": 5.0})
    net.set_masks(np.zeros(1, np.uint8), np.ones(1, np.uint8))
    kM = p["dt"] / p["tauM"]
    for n in (1, 10, 200, 1500):
        net.forward(n, seed=1, record=False)
        want = p["vRest"] + 5.0 * (1 - (1 - kM) ** n)
        assert net.state("v")[0] == pytest.approx(want, abs=2e-4), n


def test_delivery_lands_after_the_delay_with_the_lifdiff_weight():
    """A forced spike of neuron 0 at step 0 reaches neuron 1 at step D with
    counts * inScale * (sign * wSyn * outScale * exp(logGain) * res * amp)
    (src/lifdiff.js:278-285), then decays with dE."""
    net, p = _lone(2, edges=[(0, 1)], counts=[3.0], over={"wSyn": 0.4, "tRef": 2.2})
    drive = np.array([1e7, 0.0], np.float32)          # a draw always fires
    net.set_input("drive", drive)
    net.set_input("outScale", np.array([1.5, 1.0], np.float32))
    net.set_input("logGain", np.array([0.25, 0.0], np.float32))
    net.set_driven(np.array([0], np.int32))
    D = P.nslots(p) - 1
    dE = np.exp(-p["dt"] / p["tauSyn"])
    net.forward(D, seed=1, record=False)
    assert net.state("gE")[1] == 0.0
    net.forward(D + 1, seed=1, record=False)
    eff = 1.0 * 0.4 * 1.5 * np.exp(np.float64(np.float32(0.25))) * 1.0 * 1.0
    g = np.float32(np.float64(np.float32(0.0)) + 3.0 * 1.0 * eff)
    assert net.state("gE")[1] == np.float32(np.float64(g) * dE)


def test_forced_refractory_countdown_and_soft_amplitude():
    """A forced spike sets refr = tRef + dt, then refr -= dt per step; in soft
    mode the amplitude is the logistic 1 / (1 + exp(-(v - thr) / beta))."""
    net, p = _lone(1, over={"tRef": 0.5})
    net.set_input("drive", np.array([1e7], np.float32))
    net.set_driven(np.array([0], np.int32))
    net.forward(1, seed=1, record=False)
    assert net.state("refr")[0] == np.float32(np.float64(np.float32(0.5 + 0.1)) - 0.1)
    net2, p2 = _lone(1, over={"soft": True, "surrogateBeta": 1.5, "laminaBias": 30.0})
    net2.set_masks(np.zeros(1, np.uint8), np.ones(1, np.uint8))
    net2.forward(1, seed=1, record=True)
    kM = p2["dt"] / p2["tauM"]
    u = p2["vRest"] + (p2["vRest"] - p2["vRest"] + 30.0) * kM
    s = 1.0 / (1.0 + np.exp(-(u - p2["vThresh"]) / 1.5))
    assert net2.state("spikeCount")[0] == np.float32(s)
    assert net2.state("v")[0] == np.float32(s * p2["vReset"] + (1 - s) * u)


# ---------------------------------------------------------------- the drive as an input

def test_drive_gradient_matches_finite_differences(oracle_built):
    """driveIdx with driveSoft (the smooth model scripts/vis_grad_check.mjs
    checks the drive gradient against): dL/d rate at each epoch, against
    central differences of the schedule entry, with grad_check's tolerance."""
    pr = dc.Problem().with_schedule(nD=20, epoch=8, steps=STEPS, soft=True)
    idx, ep, vals, soft = pr.sched
    pr.dL[idx[::2]] = 1.0            # half of the driven neurons are scored themselves
    net = pr.oracle()
    L = net.forward(STEPS, seed=1)
    net.backward()
    gd = net.drive_grad()
    h = 5.0
    checked = 0
    for e, k in ((0, 0), (1, 4), (2, 6), (4, 10), (6, 18), (7, 5), (3, 14)):
        up, dn = vals.copy(), vals.copy()
        up[e, k] += h
        dn[e, k] -= h
        pr.sched = [idx, ep, up, soft]
        Lu = pr.oracle().forward(STEPS, seed=1, record=False)
        pr.sched = [idx, ep, dn, soft]
        Ld = pr.oracle().forward(STEPS, seed=1, record=False)
        pr.sched = [idx, ep, vals, soft]
        fd = (Lu - Ld) / (2 * h)
        ok, rel = dc.fd_ok(gd[e, k], fd, h, L)
        assert ok, (e, k, gd[e, k], fd, rel)
        checked += abs(fd) > 4e-6 * abs(L) / h * 3
    assert checked >= 4          # most rows are above the finite-difference noise floor


def test_per_epoch_loss_gradients_match_finite_differences(oracle_built):
    """dLdSpikePerEpoch (src/lifdiff.js:519-522): a loss weighting each
    neuron's spikes per epoch; the scalar and drive gradients of that loss
    against central differences."""
    pr = dc.Problem().with_schedule(nD=12, epoch=8, steps=STEPS, soft=True, seed=9)
    rng = np.random.default_rng(4)
    nEp = -(-STEPS // 8)
    pr.dper = (rng.normal(0, 1, size=(nEp, pr.g.N)).astype(np.float32), 8)
    net = pr.oracle()
    L = net.forward(STEPS, seed=1)
    g = net.backward()
    for name in ("wSyn", "vThresh", "adaptInc"):
        h = dc.H[name]
        fd = (_loss(pr, params={name: pr.p[name] + h}) - _loss(pr, params={name: pr.p[name] - h})) / (2 * h)
        ok, rel = dc.fd_ok(g["params"][name], fd, h, L)
        assert ok, (name, g["params"][name], fd, rel)
    # the loss is the per-epoch weighted count, not the flat one
    flat = pr.oracle()
    flat_L = flat.forward(STEPS, seed=1, record=False)
    assert flat_L == L
    pr.dper = None
    assert _loss(pr) != L


def test_schedule_gradients_do_not_depend_on_checkpoint_length(oracle_built):
    """Segments that start inside an epoch reload its rates (src/lifdiff.js:
    406-413): the drive gradient is the same for checkpoints of 16 and 7 steps."""
    out = []
    for K in (16, 7):
        pr = dc.Problem(over={"checkpoint": K}).with_schedule(nD=15, epoch=8, steps=STEPS, soft=True, zeros=True)
        net = pr.oracle()
        net.forward(STEPS, seed=1)
        g = net.backward()
        out.append((g["params"], net.drive_grad()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def test_hard_schedule_matches_lif_c_with_set_drive_each_epoch(oracle_built):
    """Hard mode, no synapses: the schedule is setDriveOne at each epoch start,
    so neurons set to 0 leave the driven set and come back at its end; the
    forced spikes and the stream equal lif.c driven the same way."""
    from oracle import OracleBrain, OracleDiff
    N, ep, T = 200, 25, 1000
    g = synth.Graph(N=N, indptr=np.zeros(N + 1, np.uint32), indices=np.zeros(0, np.uint32),
                    counts=np.zeros(0, np.int32), exc=np.ones(N, bool))
    rng = np.random.default_rng(12)
    idx = rng.permutation(N)[:60].astype(np.int32)
    vals = rng.uniform(20, 400, size=(T // ep, 60)).astype(np.float32)
    vals[rng.random(vals.shape) < 0.35] = 0.0
    base_drv = np.arange(0, N, 11, dtype=np.int32)
    drive = np.zeros(N, np.float32)
    drive[base_drv] = 90.0
    pd = P.diff_merged({"dt": 0.1, "tRef": 2.2})
    od = OracleDiff(N, g.indptr, g.indices, np.zeros(0, np.float32), np.ones(N, np.float32), pd)
    od.set_input("drive", drive)
    od.set_driven(base_drv)
    od.set_loss(np.zeros(N, np.float32))
    od.set_drive_schedule(idx, ep, vals)
    od.forward(T, seed=3)
    off, sidx, amp, forced = od.spikes()
    ob = OracleBrain(g, P.config_params(tRef=2.2), seed=3, W=np.zeros(0, np.float32), sign=np.ones(N, np.float32))
    for i in base_drv:
        ob.setDriveOne(int(i), 90.0)
    nfs, ras = [], []
    for e in range(T // ep):
        for k, i in enumerate(idx):
            ob.setDriveOne(int(i), float(vals[e, k]))
        nf, r, _ = ob.run(ep, raster_cap=ep * N)
        nfs += nf.tolist()
        ras += r.tolist()
    assert forced.all() and np.diff(off).tolist() == nfs
    assert sidx.tolist() == ras
    assert od.rng == ob.rng


def test_per_neuron_beta_equal_to_the_scalar_changes_nothing(base):
    """surrogateBeta as a Float32Array (src/lifdiff.js:89-95): the same value
    on every neuron gives the scalar run's loss and gradients exactly."""
    pr, L, g = base
    net = pr.oracle()
    net.set_input("surrogateBeta", np.full(pr.g.N, pr.p["surrogateBeta"], np.float32))
    L2 = net.forward(STEPS, seed=1)
    g2 = net.backward()
    assert L2 == L
    assert g2["params"] == g["params"]
    assert np.array_equal(g2["logGain"], g["logGain"])


def test_per_neuron_beta_gradients_match_finite_differences():
    """A split surrogate (a wider one on a quarter of the neurons, as the
    reference gives visual projection neurons, src/diffsetup.js:93-101) on the
    soft model: the nine scalar adjoints against central differences, and the
    array is read per neuron (a changed entry changes the run)."""
    pr = dc.Problem()
    beta = np.full(pr.g.N, pr.p["surrogateBeta"], np.float32)
    beta[: pr.g.N // 4] = 3.0
    pr.inputs["surrogateBeta"] = beta
    net = pr.oracle()
    L = net.forward(STEPS, seed=1)
    g = net.backward()
    assert L != _loss(dc.Problem())
    for name in P.DIFF_PARAMS:
        h = dc.H[name]
        fd = (_loss(pr, params={name: pr.p[name] + h}) - _loss(pr, params={name: pr.p[name] - h})) / (2 * h)
        ok, rel = dc.fd_ok(g["params"][name], fd, h, L)
        assert ok, (name, g["params"][name], fd, rel)


def test_given_in_scale_is_taken_verbatim(base):
    """opts.inScale (src/lifdiff.js:115-118, 160): the derived inScale handed
    back as an input gives the same run; gInScale is then dL/d inScale, checked
    against central differences of the given array."""
    pr, L, g = base
    derived = pr.oracle()
    derived.forward(1, seed=1, record=False)
    inS = derived.state("inScale").copy()
    net = pr.oracle(inputs={"inScale": inS})
    assert net.forward(STEPS, seed=1) == L
    g2 = net.backward()
    assert g2["params"] == g["params"]
    h = 1e-2
    for q in np.argsort(-np.abs(g2["inScale"]))[:4].tolist():
        up, dn = inS.copy(), inS.copy()
        up[q] = np.float32(float(up[q]) + h)
        dn[q] = np.float32(float(dn[q]) - h)
        fd = (_loss(pr, inputs={"inScale": up}) - _loss(pr, inputs={"inScale": dn})) / (2 * h)
        ok, rel = dc.fd_ok(float(g2["inScale"][q]), fd, h, L)
        assert ok, (q, g2["inScale"][q], fd, rel)
    # cleared, the derived inScale is back
    net.set_input("inScale", None)
    assert net.forward(STEPS, seed=1) == L


# ---------------------------------------------------------------- forward's spikeTrain

def _forced_lists(net):
    """The forced spikes of each recorded step, in list order."""
    off, idx, _, forced = net.spikes()
    return [idx[off[t]:off[t + 1]][forced[off[t]:off[t + 1]]].tolist() for t in range(off.size - 1)]


@pytest.mark.parametrize("soft", [False, True])
def test_spike_train_is_the_driven_spikes_it_replays(oracle_built, soft):
    """A train that lists, step by step, the spikes a driven run forced gives a
    run without drive the same forward pass bit for bit (state, spike lists)
    and the same gradients: forced like a driven spike, without a draw, at the
    head of the step's list (src/lifdiff.js:290-307)."""
    pr = dc.Problem(N=300, seed=17, over={"soft": soft, "tRef": 1.0}, drive_every=4, drive_hz=150.0)
    a = pr.oracle()
    T = 120
    La = a.forward(T, seed=3)
    train = _forced_lists(a)
    assert sum(len(x) for x in train) > 100
    b = pr.oracle(inputs={"drive": np.zeros(pr.g.N, np.float32)})
    b.set_driven(np.zeros(0, np.int32))
    Lb = b.forward(T, seed=99, spikeTrain=train)
    assert Lb == La
    for k in ("v", "gE", "gI", "adapt", "res", "refr", "spikeCount"):
        assert np.array_equal(a.state(k).view(np.uint32), b.state(k).view(np.uint32)), k
    oa, ia, aa, fa = a.spikes()
    ob, ib, ab, fb = b.spikes()
    assert np.array_equal(oa, ob) and np.array_equal(ia, ib) and np.array_equal(aa, ab) and np.array_equal(fa, fb)
    ga = a.backward()
    gb = b.backward()
    assert ga["params"] == gb["params"]
    assert np.array_equal(ga["logGain"], gb["logGain"]) and np.array_equal(ga["edge"], gb["edge"])


def test_spike_train_skips_refractory_and_repeated_neurons(oracle_built):
    """A listed neuron that is refractory -- listed twice in a step, listed on
    the next step inside tRef, or drawn by its drive in the same step -- is
    skipped (src/lifdiff.js:303)."""
    net, p = _lone(3, over={"tRef": 0.25})
    net.set_input("drive", np.array([1e7, 0.0, 0.0], np.float32))   # neuron 0's draw always fires
    net.set_driven(np.array([0], np.int32))
    # forced: refr = tRef + dt = 0.35 ms, down 0.1 ms a step: refractory at
    # steps 1-3, free again at step 4
    train = [[1, 1, 0, 2], [1], [], [1], [1]]
    net.forward(5, seed=1, spikeTrain=train)
    assert _forced_lists(net) == [[0, 1, 2], [], [], [], [0, 1]]
    assert net.state("spikeCount").tolist() == [2.0, 2.0, 1.0]


def test_spike_train_burst_reaches_its_target_scaled_by_log_gain(oracle_built):
    """scripts/vnc_regression.mjs:40-79: a 20 Hz train on neuron 0 of a
    0 -> 1 edge (count 10, wSyn 5) delivers conductance and fires neuron 1; a
    log-gain of log 4 on the source makes the delivered conductance 4x."""
    net, p = _lone(2, edges=[(0, 1)], counts=[10.0], over={"wSyn": 5.0, "dt": 0.5})
    train = [[0] if (t % 5 == 0 and t < 50) else None for t in range(60)]
    gmax = 0.0
    for n in range(1, 61):
        net.forward(n, seed=1, record=False, spikeTrain=train)
        gmax = max(gmax, float(net.state("gE")[1]))
    net.forward(60, seed=1, spikeTrain=train)
    off, idx, _, forced = net.spikes()
    assert gmax > 0 and int(((idx == 1) & ~forced).sum()) > 0
    one = [[0]] + [None] * 29

    def peak(lg):
        n2, _ = _lone(2, edges=[(0, 1)], counts=[10.0], over={"wSyn": 1.0, "dt": 0.5})
        n2.set_input("logGain", np.array([lg, 0.0], np.float32))
        return max(float((n2.forward(n, seed=1, record=False, spikeTrain=one), n2.state("gE")[1])[1])
                   for n in range(1, 31))

    g0, g4 = peak(0.0), peak(float(np.log(4.0)))
    assert g0 > 0 and g4 == pytest.approx(4.0 * g0, rel=1e-6)


def test_spike_train_offsets_for_the_boundary():
    """train_csr, the host side of fb_diff_set_spike_train: offsets from 0 over
    the steps, empty steps (None or []) kept, list order kept."""
    diff = importlib.import_module("fly-brain_amd.diff")
    ptr, idx = diff.train_csr([[3, 1, 3], None, [], [7]])
    assert ptr.dtype == np.int64 and idx.dtype == np.int32
    assert ptr.tolist() == [0, 3, 3, 3, 4] and idx.tolist() == [3, 1, 3, 7]
    ptr, idx = diff.train_csr(None)
    assert ptr.tolist() == [0] and idx.size == 0
