"""GPU parity of the differentiable model (fb_diff_*) against the CPU oracle
restatement of LIFDiff, and of its gradients against central finite
differences of the GPU's own soft forward pass.

Hard threshold: the forward pass is bit-identical to the oracle (tape, state,
stream, loss).  The backward pass sums the same double terms in another order
(fan-out reductions across lanes, scalar sums across workgroups), so gradients
agree to ~1e-6 relative.  Soft threshold: the logistic's exp runs on the GPU,
so the forward agrees to rounding."""
import importlib

import numpy as np
import pytest

import diffcases as dc

pytestmark = pytest.mark.gpu

diff = importlib.import_module("fly-brain_amd.diff")
synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

STEPS = 60
STATE = ("v", "gE", "gI", "adapt", "res", "refr", "spikeCount")


def gpu_of(pr, params=None, inputs=None, edge_gain=None):
    p = {**pr.p, **(params or {})}
    net = diff.LIFDiffHip(pr.g.N, pr.g.indptr, pr.g.indices, pr.counts, pr.sign, p)
    return pr.apply(net, params, inputs, edge_gain)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(1e-30, np.abs(b).max()))


def _pair(pr, steps=STEPS, seed=1):
    o = pr.oracle()
    g = gpu_of(pr)
    Lo = o.forward(steps, seed=seed)
    Lg = g.forward(steps, seed=seed)
    return o, g, Lo, Lg


@pytest.mark.parametrize("over", [{"soft": False, "tRef": 2.2}, {"soft": False, "tRef": 2.2, "checkpoint": 7},
                                  {"soft": False, "tRef": 0.5, "delay": 0.5}])
def test_hard_forward_is_bit_identical_to_oracle(over):
    pr = dc.Problem(over=over).with_readout(R=30)
    o, g, Lo, Lg = _pair(pr, steps=150)
    off_o, idx_o, amp_o, f_o = o.spikes()
    off_g, idx_g, amp_g, f_g = g.spikes()
    assert np.array_equal(off_o, off_g)
    assert np.array_equal(idx_o, idx_g) and np.array_equal(f_o, f_g)
    assert np.array_equal(amp_o, amp_g)
    for k in STATE:
        assert np.array_equal(o.state(k).view(np.uint32), g.state(k).view(np.uint32)), k
    assert o.rng == g.rng
    # the loss sums the same terms in another order
    assert Lg == pytest.approx(Lo, rel=1e-12)
    so, ro = o.readout()
    sg, rg = g.readout()
    assert np.array_equal(so, sg) and np.array_equal(ro, rg)


def test_hub_targets_with_many_arrivals_match_oracle():
    """Two targets that every neuron projects to (one row entry twice), under a
    drive that fires dozens of sources a step: their accumulators take more
    than 32 arrivals in one step, which the long pass sorts in LDS.  Tape,
    state and gradients against the oracle, forward and replay."""
    from types import SimpleNamespace
    pr = dc.Problem(N=400, drive_every=2, drive_hz=900.0, over={"soft": False, "tRef": 0.5, "minSyn": 1})
    g0 = pr.g
    rows = []
    for j in range(g0.N):
        r = list(g0.indices[g0.indptr[j]:g0.indptr[j + 1]])
        c = list(g0.counts[g0.indptr[j]:g0.indptr[j + 1]])
        rows.append((r + [1, 2, 1], c + [3.0, 2.0, 1.0]))
    indptr = np.concatenate([[0], np.cumsum([len(r) for r, _ in rows])]).astype(np.int64)
    indices = np.concatenate([r for r, _ in rows]).astype(np.int32)
    counts = np.concatenate([c for _, c in rows]).astype(np.float32)
    pr.g = SimpleNamespace(N=g0.N, indptr=indptr, indices=indices, counts=counts, exc=g0.exc, E=int(indices.size))
    pr.counts = P.gated_counts(counts, indices, pr.p["minSyn"])
    print("hub targets with many arrivals a step", flush=True)
    o, g, Lo, Lg = _pair(pr, steps=120)
    off_o, idx_o, amp_o, f_o = o.spikes()
    off_g, idx_g, amp_g, f_g = g.spikes()
    assert np.array_equal(off_o, off_g) and np.array_equal(idx_o, idx_g) and np.array_equal(f_o, f_g)
    for k in STATE:
        assert np.array_equal(o.state(k).view(np.uint32), g.state(k).view(np.uint32)), k
    per_step = np.diff(off_o)
    assert per_step.max() > 40          # every spike arrives at both hubs, one of them twice
    go, gg = o.backward(), g.backward()
    assert rel_err(gg["logGain"], go["logGain"]) < 1e-5
    assert rel_err(gg["edge"], go["edge"]) < 1e-5


@pytest.mark.parametrize("over", [{"soft": False, "tRef": 2.2}, {"soft": False, "tRef": 2.2, "adjClip": 0.05}])
def test_hard_gradients_match_oracle(over):
    pr = dc.Problem(over=over).with_readout(R=30)
    o, g, Lo, Lg = _pair(pr, steps=150)
    go, gg = o.backward(), g.backward()
    for k in P.DIFF_PARAMS:
        a, b = gg["params"][k], go["params"][k]
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b)) + 1e-6 * abs(b), (k, a, b)
    assert rel_err(gg["logGain"], go["logGain"]) < 1e-5
    assert rel_err(gg["edge"], go["edge"]) < 1e-5
    assert rel_err(gg["inScale"], go["inScale"]) < 1e-5
    assert np.array_equal(gg["clamped"], go["clamped"])


def test_soft_forward_and_gradients_match_oracle():
    pr = dc.Problem()
    o, g, Lo, Lg = _pair(pr)
    assert Lg == pytest.approx(Lo, rel=1e-6)
    for k in STATE:
        assert np.allclose(g.state(k), o.state(k), rtol=1e-5, atol=1e-5), k
    go, gg = o.backward(), g.backward()
    for k in P.DIFF_PARAMS:
        assert gg["params"][k] == pytest.approx(go["params"][k], rel=1e-4, abs=1e-8), k
    assert rel_err(gg["logGain"], go["logGain"]) < 1e-4
    assert rel_err(gg["edge"], go["edge"]) < 1e-4


def test_gpu_gradients_match_finite_differences_of_gpu_forward():
    """grad_check.mjs's comparison on the GPU alone: nine scalars and
    per-neuron log-gains of the soft model."""
    pr = dc.Problem(N=600, seed=21)
    g = gpu_of(pr)
    L = g.forward(STEPS, seed=1)
    gr = g.backward()

    def loss(**kw):
        return gpu_of(pr, **kw).forward(STEPS, seed=1, record=False)

    for k in P.DIFF_PARAMS:
        h = dc.H[k]
        fd = (loss(params={k: pr.p[k] + h}) - loss(params={k: pr.p[k] - h})) / (2 * h)
        ok, rel = dc.fd_ok(gr["params"][k], fd, h, L)
        assert ok, (k, gr["params"][k], fd, rel)
    h = 2e-2
    for i in (0, 5, 200, 400):
        up = pr.inputs["logGain"].copy()
        dn = up.copy()
        up[i] += h
        dn[i] -= h
        fd = (loss(inputs={"logGain": up}) - loss(inputs={"logGain": dn})) / (2 * h)
        ok, rel = dc.fd_ok(gr["logGain"][i], fd, h, L)
        assert ok, (i, gr["logGain"][i], fd, rel)


def test_checkpoint_length_does_not_change_gradients():
    out = []
    for K in (16, 5, 150):
        pr = dc.Problem(over={"soft": False, "tRef": 2.2, "checkpoint": K}).with_readout(R=20)
        g = gpu_of(pr)
        g.forward(150, seed=3)
        out.append(g.backward())
    for o in out[1:]:
        for k in P.DIFF_PARAMS:
            if k != "sizeAlpha":      # reads gInScale, summed with float atomics
                assert o["params"][k] == out[0]["params"][k], k
        assert np.array_equal(o["logGain"], out[0]["logGain"])
        assert np.array_equal(o["edge"], out[0]["edge"])


def test_type_gain_gradients():
    pr = dc.Problem(over={"soft": False, "tRef": 2.2})
    types = np.random.default_rng(5).integers(0, 12, size=pr.g.N).astype(np.int32)
    g = gpu_of(pr)
    g.forward(100, seed=2)
    gr = g.backward()
    tg = g.type_grads(types, 12)
    want = np.zeros(12)
    np.add.at(want, types, gr["logGain"].astype(np.float64) / pr.inputs["outScale"].astype(np.float64))
    assert np.allclose(tg, want, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- the full-scale workload

def c4_problem(soft=False, steps=2000):
    """165,122 neurons, ~8M edges, COBA (eInh -80), dt 0.1, 1% driven at
    U(100, 200) Hz, 500 neuron types with lognormal(0, 0.2) gains on outgoing
    weights, 1,000 readouts with U(0, 50) Hz targets and a 20 ms rate filter,
    checkpoints every 100 steps, fast-sigmoid beta 0.1 mV."""
    g = synth.config2_graph()
    p = P.diff_merged({"dt": 0.1, "tRef": 2.2, "delay": 1.8, "wSyn": 0.275, "eInh": -80.0,
                       "surrogateBeta": 0.1, "checkpoint": 100, "soft": soft})
    drv = synth.driven_set(g.N, 0.01, 4)
    rates = synth.drive_rates(drv.size, 100.0, 200.0, 5)
    types = synth.type_ids(g.N, 500, 6)
    gains = synth.type_gains(0)
    rng = np.random.default_rng(7)
    ro = np.sort(rng.choice(g.N, size=1000, replace=False)).astype(np.int32)
    tgt = np.random.default_rng(8).uniform(0, 50, size=1000).astype(np.float32)
    sign = np.where(g.exc, 1.0, -1.0).astype(np.float32)
    return g, p, drv, rates, types, gains, ro, tgt, sign


def c4_net(g, p, drv, rates, types, gains, ro, tgt, sign, cls, out_scale=None):
    net = cls(g.N, g.indptr, g.indices, g.counts.astype(np.float32), sign, p)
    drive = np.zeros(g.N, np.float32)
    drive[drv] = rates
    net.set_input("drive", drive)
    net.set_driven(drv)
    net.set_input("outScale", gains[types] if out_scale is None else out_scale)
    net.set_loss(np.zeros(g.N, np.float32), 0)
    net.set_readout(ro, tgt, 20.0)
    return net


def test_full_scale_forward_adjoint_matches_oracle():
    """The full-scale workload, 600 steps (the oracle's backward at this size
    takes seconds): hard-threshold tape bit-identical, gradients agree."""
    from oracle import OracleDiff
    w = c4_problem(steps=600)
    gpu = c4_net(*w, diff.LIFDiffHip)
    cpu = c4_net(*w, lambda N, ip, ix, c, s, p: OracleDiff(N, ip, ix, c, s, p))
    Lg = gpu.forward(600, seed=1)
    Lc = cpu.forward(600, seed=1)
    assert Lg == pytest.approx(Lc, rel=1e-12)
    og, oc = gpu.spikes(), cpu.spikes()
    for a, b in zip(og, oc):
        assert np.array_equal(a, b)
    gg, gc = gpu.backward(), cpu.backward()
    assert rel_err(gg["logGain"], gc["logGain"]) < 1e-5
    assert rel_err(gg["edge"], gc["edge"]) < 1e-5
    for k in ("wSyn", "vThresh", "inhGain", "eInh"):
        assert gg["params"][k] == pytest.approx(gc["params"][k], rel=1e-6), k
    types = w[4]
    tg = gpu.type_grads(types, 500)
    assert np.count_nonzero(tg) > 400


def test_full_scale_type_gain_gradients_match_finite_differences():
    """Soft model on the full-scale workload: 32 random type gains (seed 9),
    central differences with eps 1e-3 on log gain, relative error <= 1e-3
    (with grad_check's f32 noise floor)."""
    w = c4_problem(soft=True)
    g, p, drv, rates, types, gains, ro, tgt, sign = w
    net = c4_net(*w, diff.LIFDiffHip)
    L = net.forward(2000, seed=1)
    gr = net.backward(edges=False)
    tg = net.type_grads(types, 500) * gains.astype(np.float64)      # d L / d log gain_T
    pick = np.random.default_rng(9).choice(500, size=32, replace=False)
    h = 1e-3
    fails = []
    for T in pick:
        up = gains.astype(np.float64).copy()
        dn = up.copy()
        up[T] *= np.exp(h)
        dn[T] *= np.exp(-h)
        net.set_input("outScale", up.astype(np.float32)[types])
        Lp = net.forward(2000, seed=1, record=False)
        net.set_input("outScale", dn.astype(np.float32)[types])
        Lm = net.forward(2000, seed=1, record=False)
        fd = (Lp - Lm) / (2 * h)
        ok, rel = dc.fd_ok(tg[T], fd, h, L)
        if not ok:
            fails.append((int(T), float(tg[T]), fd, rel))
    assert not fails, fails


def test_lockstep_members_match_separate_runs():
    """Members of LIFDiffBatch, stepped in lockstep with their own inputs,
    seeds and windows (not multiples of the checkpoint interval), give
    exactly what each gives alone."""
    pr = dc.Problem(over={"soft": False, "tRef": 2.2}).with_readout(R=20)
    B = 4
    seeds = [3, 4, 5, 6]
    steps = [120, 150, 37, 150]
    rng = np.random.default_rng(17)
    scales = [(pr.inputs["outScale"] * np.exp(rng.normal(0, 0.05, pr.g.N))).astype(np.float32) for _ in range(B)]
    types = rng.integers(0, 9, size=pr.g.N).astype(np.int32)
    bt = diff.LIFDiffBatch(pr.g.N, pr.g.indptr, pr.g.indices, pr.counts, pr.sign, pr.p, B)
    for b in range(B):
        pr.apply(bt[b], inputs={"outScale": scales[b]})
    res = bt.run(steps, seeds, edges=True)
    total = np.zeros(9)
    for b in range(B):
        one = gpu_of(pr, inputs={"outScale": scales[b]})
        L = one.forward(steps[b], seed=seeds[b])
        g = one.backward()
        assert res[b][0] == L
        assert np.array_equal(res[b][1]["logGain"], g["logGain"])
        assert np.array_equal(res[b][1]["edge"], g["edge"])
        for k in P.DIFF_PARAMS:
            if k != "sizeAlpha":      # reads gInScale, summed with float atomics
                assert res[b][1]["params"][k] == g["params"][k], (b, k)
        total += one.type_grads(types, 9)
    assert np.allclose(bt.type_grads(types, 9), total, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- the drive as an input

def test_soft_drive_schedule_matches_oracle():
    """driveIdx with driveSoft: the loss, the scalar gradients and the drive
    gradient per epoch (LIFDiff.backward's `drive`) against the oracle."""
    pr = dc.Problem().with_schedule(nD=20, epoch=8, steps=STEPS, soft=True)
    pr.dL[pr.sched[0][::2]] = 1.0
    print("drive schedule: soft drive", flush=True)
    o, g, Lo, Lg = _pair(pr)
    assert Lg == pytest.approx(Lo, rel=1e-6)
    go, gg = o.backward(), g.backward()
    for k in P.DIFF_PARAMS:
        assert gg["params"][k] == pytest.approx(go["params"][k], rel=1e-4, abs=1e-8), k
    do, dg = o.drive_grad(), g.drive_grad()
    assert np.abs(do).max() > 0
    assert rel_err(dg, do) < 1e-4


@pytest.mark.parametrize("K", [16, 7])
def test_hard_drive_schedule_matches_oracle(K):
    """Hard threshold, rates that drop to 0 and come back (the driven set
    changes order between epochs), checkpoints that start inside an epoch: the
    tape is bit-identical to the oracle and the drive gradient agrees."""
    pr = dc.Problem(over={"soft": False, "tRef": 2.2, "checkpoint": K}).with_schedule(
        nD=30, epoch=10, steps=150, soft=False, zeros=True)
    pr.dL[pr.sched[0]] = 1.0
    print(f"drive schedule: hard, checkpoint {K}", flush=True)
    o, g, Lo, Lg = _pair(pr, steps=150)
    off_o, idx_o, amp_o, f_o = o.spikes()
    off_g, idx_g, amp_g, f_g = g.spikes()
    assert np.array_equal(off_o, off_g) and np.array_equal(idx_o, idx_g) and np.array_equal(f_o, f_g)
    assert f_o.sum() > 50
    for k in STATE:
        assert np.array_equal(o.state(k).view(np.uint32), g.state(k).view(np.uint32)), k
    assert o.rng == g.rng
    go, gg = o.backward(), g.backward()
    do, dg = o.drive_grad(), g.drive_grad()
    assert np.abs(do).max() > 0
    assert rel_err(dg, do) < 1e-5
    assert rel_err(gg["logGain"], go["logGain"]) < 1e-5


def test_per_epoch_loss_matches_oracle():
    """dLdSpikePerEpoch: the weighted loss and its gradients against the oracle."""
    pr = dc.Problem().with_schedule(nD=12, epoch=8, steps=STEPS, soft=True, seed=9)
    rng = np.random.default_rng(4)
    pr.dper = (rng.normal(0, 1, size=(-(-STEPS // 8), pr.g.N)).astype(np.float32), 8)
    print("per-epoch spike weights", flush=True)
    o, g, Lo, Lg = _pair(pr)
    assert Lg == pytest.approx(Lo, rel=1e-6)
    go, gg = o.backward(), g.backward()
    for k in P.DIFF_PARAMS:
        assert gg["params"][k] == pytest.approx(go["params"][k], rel=1e-4, abs=1e-8), k
    assert rel_err(g.drive_grad(), o.drive_grad()) < 1e-4


@pytest.mark.parametrize("published", [True, False])
def test_gpu_reproduces_the_reference_grad_check(published):
    """The reference's gradient check on its 800-neuron subgraph of the real
    connectome (scripts/grad_check.mjs 800 120, soft model): with the
    published problem the GPU adjoint reproduces the published table
    (docs/33-differentiable-brain.md:50-62) to 1e-5 relative, and on both
    problems it agrees with the oracle to 1e-5 relative (the logistic's exp
    is the device's)."""
    import gradcheck as G
    if not G.available():
        pytest.skip("grad-check subgraph fixture absent")
    print(f"grad_check subgraph on the GPU, published={published}", flush=True)
    gc = G.GradCheck(published=published)
    o = gc.oracle()
    Lo = o.forward(G.STEPS, seed=1)
    go = o.backward()
    net = gc.apply(diff.LIFDiffHip(gc.N, gc.indptr, gc.indices, gc.counts, gc.sign, gc.p))
    Lg = net.forward(G.STEPS, seed=1)
    gg = net.backward(edges=False)
    assert abs(Lg - Lo) <= 1e-5 * abs(Lo)
    for k in G.PUBLISHED:
        a, b = gg["params"][k], go["params"][k]
        assert abs(a - b) <= 1e-5 * abs(b) + 1e-9, (k, a, b)
        if published:
            pub = G.PUBLISHED[k]
            assert abs(a - pub) <= 1e-5 * abs(pub) + 5e-7, (k, a, pub)
    assert rel_err(gg["logGain"], go["logGain"]) <= 1e-5
    if published:
        for i, pub in G.PUBLISHED_LOGGAIN.items():
            assert abs(float(gg["logGain"][i]) - pub) <= 1e-5 * abs(pub) + 5e-7, (i, gg["logGain"][i], pub)
    net.destroy()


@pytest.mark.parametrize("over", [{}, {"depU": 0.0}, {"dt": 0.1, "tRef": 2.2}])
def test_exact32_forward_is_lifc_bit_for_bit(over):
    """exact32 (src/lifdiff.js:105-109) on the GPU: the tape, every state
    array and the stream equal the oracle's exact32 run, which equals the
    reference's lif.c bit for bit (tests/test_exact32.py)."""
    import exact32cases as xc
    from oracle import available
    print(f"exact32 forward on the GPU: {over}", flush=True)
    c = xc.Case(over=over)
    steps = 400
    o = c.oracle()
    o.forward(steps, seed=3)
    net = c.apply(diff.LIFDiffHip(c.g.N, c.g.indptr, c.g.indices, c.counts, c.sign, c.p))
    net.forward(steps, seed=3)
    off_o, idx_o, amp_o, f_o = o.spikes()
    off_g, idx_g, amp_g, f_g = net.spikes()
    assert np.array_equal(off_o, off_g) and np.array_equal(idx_o, idx_g) and np.array_equal(f_o, f_g)
    for k in STATE:
        assert np.array_equal(o.state(k).view(np.uint32), net.state(k).view(np.uint32)), k
    assert o.rng == net.rng
    b = c.lifc(seed=3, kind="ref" if available("ref") else "oracle")
    nf, ras, _ = b.run(steps, raster_cap=steps * c.g.N)
    assert np.array_equal(np.diff(off_g), nf) and np.array_equal(idx_g, ras)
    assert np.array_equal(net.state("v").view(np.uint32), b.array("v").view(np.uint32))
    net.destroy()


def test_exact32_refuses_a_threshold_mask():
    import exact32cases as xc
    c = xc.Case()
    net = c.apply(diff.LIFDiffHip(c.g.N, c.g.indptr, c.g.indices, c.counts, c.sign, c.p))
    m = np.zeros(c.g.N, np.uint8)
    m[3] = 1
    net.set_masks(m, c.bias_mask)
    with pytest.raises(RuntimeError, match="thrMask"):
        net.forward(5, seed=1)
    net.destroy()


def test_per_neuron_beta_and_given_in_scale_match_oracle():
    """A split surrogateBeta (src/lifdiff.js:89-95) and a caller-given
    inScale (src/lifdiff.js:115-118): GPU loss and gradients against the
    oracle, soft model, to 1e-5 relative."""
    print("per-neuron beta and given inScale on the GPU", flush=True)
    pr = dc.Problem()
    beta = np.full(pr.g.N, pr.p["surrogateBeta"], np.float32)
    beta[: pr.g.N // 4] = 3.0
    inS = (0.3 + np.random.default_rng(3).random(pr.g.N)).astype(np.float32)
    pr.inputs["surrogateBeta"] = beta
    pr.inputs["inScale"] = inS
    o, g, Lo, Lg = _pair(pr)
    go, gg = o.backward(), g.backward(edges=True)
    assert abs(Lg - Lo) <= 1e-5 * abs(Lo)
    for k in P.DIFF_PARAMS:
        assert abs(gg["params"][k] - go["params"][k]) <= 1e-5 * abs(go["params"][k]) + 1e-9, k
    assert rel_err(gg["logGain"], go["logGain"]) <= 1e-5
    assert rel_err(gg["inScale"], go["inScale"]) <= 1e-5
    assert np.array_equal(g.state("inScale"), inS)
    g.set_input("inScale", None)
    g.set_input("surrogateBeta", None)
    L0 = g.forward(STEPS, seed=1)
    assert abs(L0 - dc.Problem().oracle().forward(STEPS, seed=1)) <= 1e-5 * abs(L0)
    g.destroy()


@pytest.mark.parametrize("drive_every", [5, 0])
def test_spike_train_forward_and_gradients_match_oracle(drive_every):
    """forward(steps, {spikeTrain}) (src/lifdiff.js:298-307): a random train
    -- neurons repeated within a step, listed while refractory, listed next to
    driven neurons -- on top of the drive (or alone); the tape, state, stream
    and loss bit-identical to the oracle, the gradients (replay from the tape's
    forced spikes) to 1e-5, and the train applies to its own pass only."""
    pr = dc.Problem(N=300, seed=29, over={"soft": False, "tRef": 1.0}, drive_every=drive_every)
    rng = np.random.default_rng(4)
    T = 150
    train = [rng.integers(0, pr.g.N, size=rng.integers(0, 12)).tolist() if rng.random() < 0.7 else None
             for _ in range(T - 10)]
    train[3] = [7, 7, 8, 7]
    o, g = pr.oracle(), gpu_of(pr)
    Lo = o.forward(T, seed=5, spikeTrain=train)
    Lg = g.forward(T, seed=5, spikeTrain=train)
    off_o, idx_o, amp_o, f_o = o.spikes()
    off_g, idx_g, amp_g, f_g = g.spikes()
    assert np.array_equal(off_o, off_g) and np.array_equal(idx_o, idx_g)
    assert np.array_equal(f_o, f_g) and np.array_equal(amp_o, amp_g)
    assert int(f_g.sum()) > 300
    for k in STATE:
        assert np.array_equal(o.state(k).view(np.uint32), g.state(k).view(np.uint32)), k
    assert o.rng == g.rng
    assert Lg == pytest.approx(Lo, rel=1e-12)
    go, gg = o.backward(), g.backward()
    for k in P.DIFF_PARAMS:
        a, b = gg["params"][k], go["params"][k]
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b)) + 1e-6 * abs(b), (k, a, b)
    assert rel_err(gg["logGain"], go["logGain"]) < 1e-5
    assert rel_err(gg["edge"], go["edge"]) < 1e-5
    # the next pass without a train is the plain forward
    Lo2, Lg2 = o.forward(T, seed=5), g.forward(T, seed=5)
    assert np.array_equal(o.spikes()[1], g.spikes()[1]) and Lg2 == pytest.approx(Lo2, rel=1e-12)
