"""The batched forward + adjoint at its full size: 32 instances of the
full-scale adjoint workload (165,122 neurons, ~8.0M edges, COBA, hard
threshold with the fast-sigmoid surrogate) over one graph, each with its own
type-gain perturbation, drive seed and window of 500-2,000 steps, stepped in
lockstep by LIFDiffBatch.  Every instance equals the same instance run on its
own (LIFDiffHip): loss, tape, log-gain and per-edge gradients bit for bit,
and the batch's summed type-gain gradient equals the sum of the separate
runs'."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

diff = importlib.import_module("fly-brain_amd.diff")
synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")

B = 32


def workload():
    """The workload of the batched adjoint: the full-scale synthetic graph,
    COBA (eInh -80 mV), dt 0.1 ms, 1% driven at U(100, 200) Hz, 500 neuron
    types with lognormal(0, 0.2) gains, 1,000 readouts (targets U(0, 50) Hz,
    20 ms filter), beta 0.1 mV, the adjoint bounded at +-10 per step."""
    g = synth.config2_graph()
    p = P.diff_merged({"dt": 0.1, "tRef": 2.2, "delay": 1.8, "wSyn": 0.275, "eInh": -80.0,
                       "surrogateBeta": 0.1, "soft": False, "adjClip": 10.0})
    drv = synth.driven_set(g.N, 0.01, 4)
    drive = np.zeros(g.N, np.float32)
    drive[drv] = synth.drive_rates(drv.size, 100.0, 200.0, 5)
    types = synth.type_ids(g.N, 500, 6)
    gains0 = synth.type_gains(0)
    ro = np.sort(np.random.default_rng(7).choice(g.N, size=1000, replace=False)).astype(np.int32)
    tgt = np.random.default_rng(8).uniform(0, 50, size=1000).astype(np.float32)
    sign = np.where(g.exc, 1.0, -1.0).astype(np.float32)
    windows = [int(np.random.default_rng(600 + b).integers(500, 2001)) for b in range(B)]
    gains = [(gains0.astype(np.float64) * np.exp(np.random.default_rng(400 + b).normal(0, 0.05, gains0.size)))
             .astype(np.float32) for b in range(B)]
    return g, p, sign, drv, drive, types, ro, tgt, windows, gains


def setup(net, drv, drive, ro, tgt, outScale):
    net.set_input("drive", drive)
    net.set_driven(drv)
    net.set_input("outScale", outScale)
    net.set_loss(np.zeros(net.N, np.float32), 0)
    net.set_readout(ro, tgt, 20.0)
    net.set_tape_capacity(4096)
    return net


def test_batched_adjoint_at_full_size_equals_separate_runs():
    g, p, sign, drv, drive, types, ro, tgt, windows, gains = workload()
    K = diff.autotune_checkpoint(g.N, max(windows), B)
    pk = {**p, "checkpoint": K}
    seeds = [500 + b for b in range(B)]
    counts = g.counts.astype(np.float32)
    print(f"C5 batch: N={g.N} E={g.E} B={B} K={K} windows {min(windows)}..{max(windows)}", flush=True)
    bt = diff.LIFDiffBatch(g.N, g.indptr, g.indices, counts, sign, pk, B)
    for b in range(B):
        setup(bt[b], drv, drive, ro, tgt, gains[b][types])
    res = bt.run(windows, seeds, edges=True)
    tape = [bt[b].spikes() for b in range(B)]
    tg_batch = bt.type_grads(types, 500)
    bt.destroy()
    print("C5 batch done; separate runs", flush=True)
    total = np.zeros(500)
    for b in range(B):
        one = setup(diff.LIFDiffHip(g.N, g.indptr, g.indices, counts, sign, pk), drv, drive, ro, tgt,
                    gains[b][types])
        L = one.forward(windows[b], seed=seeds[b])
        gr = one.backward(edges=True)
        off, idx, amp, forced = one.spikes()
        assert res[b][0] == L, b
        assert np.array_equal(tape[b][0], off) and np.array_equal(tape[b][1], idx), b
        assert np.array_equal(res[b][1]["logGain"], gr["logGain"]), b
        assert np.array_equal(res[b][1]["edge"], gr["edge"]), b
        for k in P.DIFF_PARAMS:
            if k != "sizeAlpha":      # reads gInScale, summed with float atomics
                assert res[b][1]["params"][k] == gr["params"][k], (b, k)
        assert not res[b][1]["clamped"].any()
        total += one.type_grads(types, 500)
        one.destroy()
    assert np.allclose(tg_batch, total, rtol=1e-12, atol=0)
    assert np.abs(tg_batch).max() > 0
