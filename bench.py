#!/usr/bin/env python3
"""Synaptic events/s of the connectome LIF step on one MI355X.

Workload: the full-scale synthetic connectome -- 165,122 neurons, ~8.0M
directed edges (heavy-tailed counts, ~104M synapses, hub rows of 20,000),
CUBA, dt 0.1 ms, 1% of neurons Poisson-driven at U(100, 200) Hz.  A "step" is
one 0.1 ms LIF step of the whole network (deliver, forced spikes, membrane,
threshold); K steps are timed with HIP events on the engine's stream, after W
warm-up steps, with the state already resident in HBM.

value = edge events delivered in the timed region / seconds, where an edge
event is one (arriving spike, outgoing edge) pair -- the reference's
"edges/ms" (scripts/profile_fly.mjs:19,25 of Lulzx/fly-brain).

cpu_baseline: the reference's own lif.c compiled natively (oracle/_ref) when
present, else the oracle restatement, on one host core, over a bounded sample
of the same workload.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pkg(name):
    return importlib.import_module("fly-brain_amd." + name)


def algorithmic_bytes(N: int, edge_events: int, spikes: int, steps: int, live_adapt_res: bool = False) -> int:
    """Per-step forward bytes, summed: 32 N (v, gE, gI, refr read and written)
    + 8 per edge event (u32 target + f32 weight) + 4 per emitted spike (+16 N
    when adaptation or depression are live)."""
    per_neuron = 32 + (16 if live_adapt_res else 0)
    return per_neuron * N * steps + 8 * edge_events + 4 * spikes


def cpu_env() -> dict:
    """The host the CPU legs run on: CPU model, logical CPUs, and the CPUs
    this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else []
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": len(aff) or os.cpu_count()}


class pinned:
    """Pins the calling thread to one CPU (the first it may use) for a 1-core
    baseline, and restores its mask afterwards."""

    def __enter__(self):
        self.old = os.sched_getaffinity(0) if hasattr(os, "sched_setaffinity") else None
        self.cpu = min(self.old) if self.old else None
        if self.old:
            os.sched_setaffinity(0, {self.cpu})
        return self

    def __exit__(self, *exc):
        if self.old:
            os.sched_setaffinity(0, self.old)
        return False


def cpu_baseline(g, p, drv, rates, steps: int, warm: int, repeats: int = 3):
    """lif.c (or its restatement) on one pinned core: `warm` untimed steps to
    the same steady state the GPU is timed in, then `repeats` timed samples of
    `steps` steps; the median sample is reported."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    kind = "ref" if orc.available("ref") else "oracle"
    if not orc.available(kind):
        return None
    b = orc.OracleBrain(g, p, seed=1, kind=kind)
    for i, r in zip(drv.tolist(), rates.tolist()):
        b.setDriveOne(i, r)
    outdeg = g.outdeg
    D = b.nslots - 1
    nf_all, ras_all = [], []
    nf_w, ras_w, _ = b.run(warm, raster_cap=warm * 4096)
    nf_all.append(nf_w)
    ras_all.append(ras_w)
    samples = []
    with pinned() as pin:
        for _ in range(repeats):
            t0 = time.perf_counter()
            nf, ras, _ = b.run(steps, raster_cap=steps * 4096)
            samples.append(time.perf_counter() - t0)
            nf_all.append(nf)
            ras_all.append(ras)
    nf_cat = np.concatenate(nf_all)
    ras_cat = np.concatenate(ras_all)
    start = np.concatenate([[0], np.cumsum(nf_cat)])
    order = np.argsort(samples)
    k = int(order[len(order) // 2])                       # the median sample
    lo, hi = warm + k * steps - D, warm + (k + 1) * steps - D
    events = int(outdeg[ras_cat[start[lo]:start[hi]]].sum())
    dt = samples[k]
    return {
        "value": events / dt, "unit": "synaptic events/s", "cores": 1,
        "kind": "reference" if kind == "ref" else "port",
        "sample": f"median of {repeats} samples of {steps} steps each, after {warm} untimed steps to steady "
                  "state, of the same 165,122-neuron workload; "
                  f"{'lif.c compiled natively (gcc -O2 -ffp-contract=off)' if kind == 'ref' else 'oracle/lif_oracle.c'}, "
                  f"1 thread pinned to CPU {pin.cpu}",
        "sim_s_per_wall_s": steps * p["dt"] / 1000.0 / dt,
        "ms_per_step": dt * 1000.0 / steps,
        "samples_ms_per_step": [x * 1000.0 / steps for x in samples],
        **cpu_env(),
    }


def cpu_batch_baseline(g, p, drv, type_of, gains, silenced, steps: int, members: int):
    """`members` independent instances of lif.c (or its restatement) on all
    host cores (one per thread), `steps` steps each: the one-brain-per-worker
    ensemble the reference runs (src/brainsetup.js:8)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import concurrent.futures as cf
    import numpy as np
    import oracle as orc
    kind = "ref" if orc.available("ref") else "oracle"
    pre = np.repeat(np.arange(g.N), np.diff(g.indptr.astype(np.int64)))
    outdeg = g.outdeg
    brains = []
    for b in range(members):
        W = (g.counts.astype(np.float64) * gains[b][type_of[pre]].astype(np.float64)).astype(np.float32)
        o = orc.OracleBrain(g, p, seed=b + 1, kind=kind, W=W, sign=g.sign)
        o.setDrive(drv.tolist(), 150.0)
        for i in silenced[b]:
            o.setThr(int(i), 1e6)
        o._sync_driven()
        brains.append(o)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(lambda o: o.run(steps, raster_cap=steps * 1024), brains))
    dt = time.perf_counter() - t0
    D = brains[0].nslots - 1
    events = 0
    for nf, ras, _ in res:
        start = np.concatenate([[0], np.cumsum(nf)])
        events += int(outdeg[ras[: start[max(0, steps - D)]]].sum())
    return {"value": events / dt, "unit": "synaptic events/s", "cores": cores,
            "kind": "reference" if kind == "ref" else "port",
            "sample": f"{members} members x {steps} steps of the 64-member ensemble workload, one member per thread",
            **cpu_env()}


def bench_batch64(args, synth, P, lif, batch, hiprt, stream):
    """The 64-member ensemble on the 16,384-neuron graph, and the plain single
    brain on the same graph for the ratio."""
    import numpy as np
    g = synth.config1_graph()
    p = P.config_params()
    drv = synth.driven_set(g.N, 0.02, 3)
    B = 64
    type_of = synth.type_ids(g.N, 500, 6)
    gains = np.stack([synth.type_gains(b) for b in range(B)])
    silenced = [synth.silenced(b, g.N) for b in range(B)]
    steps = np.array([synth.run_length(b) for b in range(B)])
    hg = lif.HipGraph.from_synth(g)

    def ensemble():
        e = batch.LIFBatch(hg, p, [b + 1 for b in range(B)], type_of, gains)
        e.setDrive(drv, 150.0)
        for b in range(B):
            e.silence(b, silenced[b])
        e.set_stream(stream.handle)
        return e
    # untimed: the autotuner measures the lookahead engine's block shapes on a
    # first ensemble of this graph and size; the timed one takes its choice
    tuner = ensemble()
    tuner.run(400)
    tuning = {"shape_neurons_per_block": tuner.stat("lookahead_shape"),
              "us_per_epoch": {str(k): tuner.stat(f"autotune_us_per_epoch_{k}") for k in (72, 80, 96, 128)},
              "steps": 400}
    tuner.destroy()
    eb = ensemble()
    eb.set_steps(steps)
    c0 = eb.counters()
    e0, e1 = hiprt.Event(), hiprt.Event()
    T = int(steps.max())
    e0.record(stream)
    eb.run(T)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    c1 = eb.counters()
    ev = c1["edge_events"] - c0["edge_events"]
    la = {"launches": eb.stat("lookahead_launches"), "kernel_ms": eb.stat("lookahead_kernel_ms"),
          "steps": eb.stat("lookahead_steps"), "fallback_steps": eb.stat("fallback_steps"),
          "shape_neurons_per_block": eb.stat("lookahead_shape"), "autotune": tuning}
    out = {"value": ev / (ms / 1000.0), "unit": "synaptic events/s (aggregate over members)", "members": B,
           "engine": "lookahead" if la["steps"] else "per-step", "lookahead": la,
           "steps": T, "member_steps": int(steps.sum()), "ms_per_step": ms / T,
           "spikes": c1["spikes"] - c0["spikes"],
           "workload": "16,384 neurons, ~1.07M edges, 64 members: own seeds, 500 per-type gains lognormal(0,0.2), "
                       "0-5% silenced, run lengths U{2000..10000}"}
    eb.destroy()
    # the plain single brain on the same graph and drive
    one = lif.LIFHip(hg, p, seed=1)
    one.setDrive(drv.tolist(), 150.0)
    one.set_stream(stream.handle)
    one.advance(200)
    c0 = one.counters()
    e0.record(stream)
    one.run_async(10000)
    e1.record(stream)
    e1.synchronize()
    ms1 = e0.elapsed_ms(e1)
    c1 = one.counters()
    plain = (c1["edge_events"] - c0["edge_events"]) / (ms1 / 1000.0)
    out["plain_single_value"] = plain
    out["plain_single_ms_per_step"] = ms1 / 10000
    out["vs_plain_single"] = out["value"] / plain
    one.destroy()
    hg.release()
    for h in (e0, e1):
        h.destroy()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_batch_baseline(g, p, drv, type_of, gains, silenced, min(args.cpu_steps, 2000), B)
    return out


def traffic_per_launch(steps: int):
    """HBM bytes of one k_persist launch of `steps` steps, from the committed
    counter summary (a line `traffic_bytes_per_launch_<steps>_steps: B`), or
    None when no launch of that length was counted."""
    path = os.path.join(ROOT, "profiles", "k_persist_hbm_counters.md")
    key = f"traffic_bytes_per_launch_{steps}_steps:"
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def adjoint_workload(synth, P):
    """The full-scale graph, COBA (eInh -80), dt 0.1 ms, 1% driven at
    U(100, 200) Hz, 500 neuron types with lognormal(0, 0.2) gains on outgoing
    weights, 1,000 readouts (targets U(0, 50) Hz, 20 ms rate filter, mean
    squared rate error), hard threshold with the fast-sigmoid surrogate
    (beta 0.1 mV), checkpoints every 100 steps.  The adjoint state is bounded
    at +-10 per step (LIFDiff's adjClip): with a 0.1 mV surrogate the
    unbounded adjoint grows past 1e35 within 2,000 steps."""
    import numpy as np
    g = synth.config2_graph()
    p = P.diff_merged({"dt": 0.1, "tRef": 2.2, "delay": 1.8, "wSyn": 0.275, "eInh": -80.0,
                       "surrogateBeta": 0.1, "checkpoint": 100, "soft": False, "adjClip": 10.0})
    drv = synth.driven_set(g.N, 0.01, 4)
    drive = np.zeros(g.N, np.float32)
    drive[drv] = synth.drive_rates(drv.size, 100.0, 200.0, 5)
    types = synth.type_ids(g.N, 500, 6)
    gains = synth.type_gains(0)
    ro = np.sort(np.random.default_rng(7).choice(g.N, size=1000, replace=False)).astype(np.int32)
    tgt = np.random.default_rng(8).uniform(0, 50, size=1000).astype(np.float32)
    sign = np.where(g.exc, 1.0, -1.0).astype(np.float32)

    def setup(net):
        net.set_input("drive", drive)
        net.set_driven(drv)
        net.set_input("outScale", gains[types])
        net.set_loss(np.zeros(g.N, np.float32), 0)
        net.set_readout(ro, tgt, 20.0)
        return net
    return g, p, sign, types, setup


def bench_adjoint(args, synth, P, diff, hiprt, stream):
    import numpy as np
    g, p, sign, types, setup = adjoint_workload(synth, P)
    T = args.adjoint_steps
    net = setup(diff.LIFDiffHip(g.N, g.indptr, g.indices, g.counts.astype(np.float32), sign, p))
    net.set_stream(stream.handle)
    net.forward(200, seed=1)              # warm-up (allocations, first launches)
    net.backward(edges=True)
    ev0 = net.stat("edge_events")
    hiprt.device_synchronize()
    w0 = time.perf_counter()
    L = net.forward(T, seed=1)
    gr = net.backward(edges=True)
    wall = time.perf_counter() - w0
    ev = net.stat("edge_events") - ev0
    fwd_ms, bwd_ms = net.stat("forward_ms"), net.stat("backward_ms")
    tg = net.type_grads(types, 500)
    secs = (fwd_ms + bwd_ms) / 1000.0
    out = {"value": ev / secs, "unit": "synaptic events/s (forward events over forward + adjoint time)",
           "steps": T, "forward_ms": fwd_ms, "backward_ms": bwd_ms, "ms_per_step": (fwd_ms + bwd_ms) / T,
           "wall_s_host_clock": wall, "loss": L, "edge_events": ev,
           "tape_spikes": net.stat("tape_entries"),
           "grad_norms": {"type_gains": float(np.linalg.norm(tg)),
                          "edges": float(np.linalg.norm(gr["edge"].astype(np.float64)))},
           "non_finite_zeroed": gr["clamped"].tolist(),
           "workload": "165,122 neurons, ~8.0M edges, COBA, 2,000 steps, 1,000 readouts (filtered-rate MSE), "
                       "gradients w.r.t. 500 type gains and every edge weight, checkpoints every 100 steps"}
    net.destroy()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        n = min(T, args.adjoint_cpu_steps)
        cpu = setup(orc.OracleDiff(g.N, g.indptr, g.indices, g.counts.astype(np.float32), sign, p))
        c0 = time.perf_counter()
        cpu.forward(n, seed=1)
        cpu.backward()
        cdt = time.perf_counter() - c0
        off, idx, amp, forced = cpu.spikes()
        D = P.nslots(p) - 1
        cev = int(g.outdeg[idx[: off[max(0, n - D)]]].sum())
        out["cpu_baseline"] = {"value": cev / cdt, "unit": "synaptic events/s", "cores": 1, "kind": "port",
                               "sample": f"forward + adjoint over the first {n} steps of the same workload, "
                                         "the C restatement of LIFDiff (gcc -O2), 1 thread",
                               "ms_per_step": cdt * 1000.0 / n}
    return out


def bench_adjoint_batch(args, synth, P, diff, hiprt):
    """32 instances of the forward + adjoint workload: instance b has its own
    type-gain perturbation (x exp(N(0, 0.05^2)) per type, seed 400+b), drive
    stream seed 500+b and window U{500..2000} steps (seed 600+b); the
    checkpoint interval comes from the memory / recompute trade-off; the
    instances run in lockstep (one launch per step kernel for all of them) and
    their type-gain gradients are summed."""
    import numpy as np
    g, p, sign, types, setup = adjoint_workload(synth, P)
    B = 32
    gains0 = synth.type_gains(0)
    windows = [int(np.random.default_rng(600 + b).integers(500, 2001)) for b in range(B)]
    K = diff.autotune_checkpoint(g.N, max(windows), B)
    bt = diff.LIFDiffBatch(g.N, g.indptr, g.indices, g.counts.astype(np.float32), sign, {**p, "checkpoint": K}, B)
    for b in range(B):
        m = bt[b]
        setup(m)
        gb = (gains0.astype(np.float64) * np.exp(np.random.default_rng(400 + b).normal(0, 0.05, gains0.size)))
        m.set_input("outScale", gb.astype(np.float32)[types])
        m.set_tape_capacity(4096)
    seeds = [500 + b for b in range(B)]
    bt.run([100] * B, seeds)                 # warm-up: allocations and first launches
    ev0 = sum(m.stat("edge_events") for m in bt.members)
    hiprt.device_synchronize()
    w0 = time.perf_counter()
    res = bt.run(windows, seeds)
    tg = bt.type_grads(types, 500)
    wall = time.perf_counter() - w0
    ev = sum(m.stat("edge_events") for m in bt.members) - ev0
    fwd_ms, bwd_ms = bt.stat("forward_ms"), bt.stat("backward_ms")
    out = {"value": ev / wall, "unit": "synaptic events/s (all instances, forward events over wall time of "
                                      "forward + adjoint + gradient reduction)",
           "instances": B, "checkpoint_interval": K, "member_steps": int(sum(windows)),
           "wall_s": wall, "forward_ms": fwd_ms, "backward_ms": bwd_ms,
           "ms_per_member_step": wall * 1000.0 / sum(windows),
           "losses_mean": float(np.mean([r[0] for r in res])),
           "type_grad_norm": float(np.linalg.norm(tg)),
           "non_finite_zeroed": int(sum(int(r[1]["clamped"].sum()) for r in res)),
           "workload": "32 instances of the forward + adjoint workload with their own gain perturbations, "
                       "drive seeds and windows of 500-2,000 steps, type-gain gradients reduced over instances"}
    bt.destroy()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_adjoint_batch_baseline(args, g, p, sign, types, setup, gains0, K, windows, seeds)
    return out


def cpu_adjoint_batch_baseline(args, g, p, sign, types, setup, gains0, K, windows, seeds):
    """The batched forward + adjoint on the host: the C restatement of LIFDiff,
    one instance per host core (the reference's one-network-per-worker), each
    instance set up as the GPU's instance b, over a bounded window."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import concurrent.futures as cf
    import numpy as np
    import oracle as orc
    P = pkg("params")
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    n_inst = min(cores, len(windows))
    steps = [min(w, args.adjoint_batch_cpu_steps) for w in windows[:n_inst]]
    insts = []
    for b in range(n_inst):
        o = setup(orc.OracleDiff(g.N, g.indptr, g.indices, g.counts.astype(np.float32), sign, {**p, "checkpoint": K}))
        gb = (gains0.astype(np.float64) * np.exp(np.random.default_rng(400 + b).normal(0, 0.05, gains0.size)))
        o.set_input("outScale", gb.astype(np.float32)[types])
        insts.append(o)

    def one(b):
        insts[b].forward(steps[b], seed=seeds[b])
        insts[b].backward()
        return b

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=n_inst) as ex:
        list(ex.map(one, range(n_inst)))
    dt = time.perf_counter() - t0
    D = P.nslots(p) - 1
    ev = 0
    for b, o in enumerate(insts):
        off, idx, amp, forced = o.spikes()
        ev += int(g.outdeg[idx[: off[max(0, steps[b] - D)]]].sum())
    return {"value": ev / dt, "unit": "synaptic events/s (all instances)", "cores": n_inst, "kind": "port",
            "sample": f"instances 0-{n_inst - 1} of the batched workload, each forward + adjoint over its first "
                      f"{args.adjoint_batch_cpu_steps} steps (or its window when shorter), one instance per thread, "
                      "the C restatement of LIFDiff (gcc -O2)",
            "member_steps": int(sum(steps)), "wall_s": dt, **cpu_env()}


def bench_cord(args, lif, hiprt, stream):
    """The reference's own ventral-cord brain (21,843 cells, 1.6 M entries,
    calibrated: COBA, dt 0.5 ms, adaptation on), built from its data files
    (tests/golden/make_cord_fixture.py).  The first 1,000 steps must give the
    reference kernel's published total (139,582 spikes); then the steady-state
    rate of the lookahead engine.  Skipped when the fixture is absent."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "_cord", "cord_brains.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    p = json.loads(str(z["params"]))
    N = int(z["indptr"].size - 1)
    g = lif.HipGraph(N, z["indptr"], z["indices"], z["W"], z["sign"])
    b = lif.LIFHip(g, p, seed=1, engine="persistent")
    for i, r in zip(z["driven"].tolist(), z["rates"].tolist()):
        b.setDriveOne(i, r)
    b.set_stream(stream.handle)
    b.advance(1000)
    first = int(b.spikeCount.sum())
    steps = max(1000, args.steps)
    c0 = b.counters()
    e0, e1 = hiprt.Event(), hiprt.Event()
    e0.record(stream)
    b.run_async(steps)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    c1 = b.counters()
    ev = c1["edge_events"] - c0["edge_events"]
    out = {"workload": "the reference's ventral-cord brain from its data files: 21,843 cells, 1,615,182 entries, "
                       "calibrated parameters (COBA, dt 0.5 ms, 4-step epochs), 60 Hz to descending and 20 Hz "
                       "to proprioceptive neurons",
           "first_1000_steps_spikes": first, "reference_total": 139582, "matches_reference": first == 139582,
           "steps": steps, "ms_per_step": ms / steps, "edge_events_per_s": ev / (ms / 1e3),
           "sim_s_per_wall_s": steps * p["dt"] / 1e3 / (ms / 1e3),
           "spikes_per_step": (c1["spikes"] - c0["spikes"]) / steps}
    b.destroy()
    g.release()
    for h in (e0, e1):
        h.destroy()
    return out


def bench_brain(args, lif, hiprt, stream):
    """The reference's whole calibrated brain (165,122 neurons, 10.5 M entries,
    COBA, dt 0.5 ms, adaptation on), built from its data files
    (tests/golden/make_brain_fixture.py), on the benchmark's sugar trace:
    olfactory receptors at 6 Hz, sugar neurons at 100 Hz for 400 ms then off
    for 300 ms, seeds 1-4.  The mean total must equal the shipped kernel's
    published 448,344 spikes.  Timed: the four 1,400-step runs.  Skipped when
    the fixture is absent."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "_brain", "calibrated_brain.npz")
    if not os.path.exists(path):
        return None
    b = pkg("connectome").load_compact(path)
    p = b["params"]
    g = lif.HipGraph(b["N"], b["indptr"], b["indices"], b["W"], b["sign"])
    lam = np.nonzero(b["bias"])[0]
    totals, ms, ev, ev1, engines = [], 0.0, 0, 0, set()
    for seed in (1, 2, 3, 4):
        net = lif.LIFHip(g, p, seed=seed, engine=args.engine)
        net.set_array("thr", b["thr"])
        net.setBias(lam, float(b["bias"][lam[0]]))
        net.setDrive(b["orn"], 6.0)
        for i in b["sugar"].tolist():
            net.setDriveOne(i, 100.0)
        net.set_stream(stream.handle)
        c0 = net.counters()
        e0, e1, e2, e3 = hiprt.Event(), hiprt.Event(), hiprt.Event(), hiprt.Event()
        e0.record(stream)
        net.run_async(800)
        e1.record(stream)
        e1.synchronize()
        for i in b["sugar"].tolist():
            net.setDriveOne(i, 0.0)
        e2.record(stream)
        net.run_async(600)
        e3.record(stream)
        e3.synchronize()
        ms += e0.elapsed_ms(e1) + e2.elapsed_ms(e3)
        c1 = net.counters()
        ev += c1["edge_events"] - c0["edge_events"]
        if seed == 1:
            ev1 = c1["edge_events"] - c0["edge_events"]
        totals.append(int(net.spikeCount.sum()))
        engines.add("persistent" if net.stat("persistent_launches") > 0 else "per-step")
        net.destroy()
        for h in (e0, e1, e2, e3):
            h.destroy()
    g.release()
    steps = 4 * 1400
    out = {"workload": "the reference's whole calibrated brain from its data files: 165,122 neurons, 10,511,038 "
                       "entries (4.2 M non-zero), COBA, dt 0.5 ms; sugar trace (3,044 olfactory receptors at 6 Hz, "
                       "34 sugar neurons at 100 Hz for 800 steps, then off for 600), seeds 1-4",
           "engine": sorted(engines), "totals": totals, "mean_total": sum(totals) / 4,
           "reference_mean_total": 448344.0, "matches_reference": sum(totals) / 4 == 448344.0,
           "steps": steps, "ms_per_step": ms / steps, "edge_events_per_s": ev / (ms / 1e3),
           "sim_s_per_wall_s": steps * p["dt"] / 1e3 / (ms / 1e3)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        kind = "ref" if orc.available("ref") else "oracle"

        class _G:
            pass
        gg = _G()
        gg.N, gg.indptr, gg.indices = b["N"], b["indptr"], b["indices"]
        o = orc.OracleBrain(gg, p, seed=1, kind=kind, W=b["W"], sign=b["sign"])
        for i in np.nonzero(b["thr"])[0].tolist():
            o.setThr(i, float(b["thr"][i]))
        o.setBias(lam, float(b["bias"][lam[0]]))
        o.setDrive(b["orn"].tolist(), 6.0)
        for i in b["sugar"].tolist():
            o.setDriveOne(i, 100.0)
        o._sync_driven()
        t0 = time.perf_counter()
        o.run(800, raster_cap=1)
        t1 = time.perf_counter()
        for i in b["sugar"].tolist():
            o.setDriveOne(i, 0.0)
        o._sync_driven()
        t2 = time.perf_counter()
        o.run(600, raster_cap=1)
        dt = (t1 - t0) + (time.perf_counter() - t2)
        out["cpu_baseline"] = {"value": ev1 / dt, "unit": "synaptic events/s", "cores": 1,
                               "kind": "reference" if kind == "ref" else "port",
                               "sample": "seed 1 of the same trace (1,400 steps), "
                                         + ("lif.c compiled natively (gcc -O2 -ffp-contract=off)" if kind == "ref"
                                            else "oracle/lif_oracle.c") + ", 1 thread",
                               "ms_per_step": dt * 1e3 / 1400, "total_spikes": int(o.spikeCount.sum()),
                               "same_total_as_gpu": int(o.spikeCount.sum()) == totals[0]}
        out["gpu_over_cpu"] = out["cpu_baseline"]["ms_per_step"] / out["ms_per_step"]
    return out


def bench_flyvis(args):
    """The reference's flyvis optic lobe (45,669 nodes, 1,513,231 edges per
    eye, two eyes on one model) from its data files: untaped steps, taped
    steps with a luminance frame each, and the reverse sweep over the tape
    into one gradient for both eyes (src/visdiff.js:170-238 drives it so).
    Host-clock times through the C ABI (uploads and copies included; the
    kernels' own durations are in the kernel trace).  Skipped when the
    fixture is absent."""
    import numpy as np
    vdir = os.path.join(ROOT, "tests", "golden", "_vision")
    if not os.path.exists(os.path.join(vdir, "flyvis.bin")):
        return None
    fv = pkg("flyvis")
    m = fv.parse_flyvis(vdir)
    eyes = [fv.FlyVisDiffHip(m)]
    eyes.append(fv.FlyVisDiffHip(shared=eyes[0].sharedParts))
    rng = np.random.default_rng(1)
    nCol = eyes[0].nCol
    grey = np.full(nCol, 0.5, np.float32)
    for e in eyes:
        e.settle(grey, 150)
    _ = [e.v for e in eyes]
    n_fwd = 2000
    t0 = time.perf_counter()
    for e in eyes:
        e.step(n_fwd)
    _ = [e.v for e in eyes]
    fwd = (time.perf_counter() - t0) / (2 * n_fwd)
    T = 200
    lums = rng.uniform(0, 1, (T, nCol)).astype(np.float32)
    t0 = time.perf_counter()
    for e in eyes:
        e.startTape(T)
        for t in range(T):
            e.setInput(lums[t])
            e.stepTaped()
    _ = [e.v for e in eyes]
    taped = (time.perf_counter() - t0) / (2 * T)
    dLdV = []
    for _ in range(2):
        d = np.zeros((T, m["N"]), np.float32)
        d[:, rng.choice(m["N"], 2000, replace=False)] = rng.normal(0, 1, (T, 2000))
        dLdV.append(d)
    g = None
    t0 = time.perf_counter()
    for k, e in enumerate(eyes):
        g = e.backward(dLdV[k], g, np.zeros(T * nCol, np.float32))
    back = (time.perf_counter() - t0) / (2 * T)
    for e in eyes[::-1]:
        e.destroy()
    out = {"workload": "the reference's flyvis optic lobe from its data files: 45,669 nodes, 1,513,231 edges per "
                       "eye, two eyes on one model; 2,000 untaped steps per eye, then 200 taped steps with a "
                       "random luminance frame each and the reverse sweep over them (loss on 2,000 nodes per step)",
           "step_us": fwd * 1e6, "taped_step_us": taped * 1e6, "backward_us_per_step": back * 1e6,
           "edges_per_s_forward": m["E"] / fwd}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        o = orc.FlyVisDiffOracle(m)
        o.setInput(grey)
        t0 = time.perf_counter()
        for _ in range(20):
            o.step()
        cf = (time.perf_counter() - t0) / 20
        o.start_tape()
        for t in range(20):
            o.setInput(lums[t])
            o.step()
        t0 = time.perf_counter()
        o.backward(dLdV[0][:20])
        cb = (time.perf_counter() - t0) / 20
        out["cpu_baseline"] = {"value": m["E"] / cf, "unit": "edges/s (forward)", "cores": 1, "kind": "port",
                               "sample": "one eye, 20 steps forward and the reverse sweep over 20 taped steps, "
                                         "oracle/lif_oracle.c (fbo_fv_step, fbo_fv_back_step), 1 thread",
                               "step_us": cf * 1e6, "backward_us_per_step": cb * 1e6}
        out["gpu_over_cpu_step"] = cf / fwd
        out["gpu_over_cpu_backward"] = cb / back
    return out


def bench_visual(args):
    """The optic lobe and the whole CNS as one differentiable model
    (src/visdiff.js, as scripts/vis_fit.mjs runs it): both eyes of the
    reference's flyvis model, its 62,157 eye-to-CNS pairs, the calibrated
    brain as the CNS (hard threshold, olfactory receptors at 6 Hz), 400 steps
    of grey then a looming disc, the loss on spikes after onset; forward and
    the adjoint through the CNS, the coupling and both eyes.  Skipped when the
    fixtures are absent."""
    import math
    import numpy as np
    vdir = os.path.join(ROOT, "tests", "golden", "_vision")
    bpath = os.path.join(ROOT, "tests", "golden", "_brain", "calibrated_brain.npz")
    if not (os.path.exists(os.path.join(vdir, "flyvis_map.json")) and os.path.exists(bpath)):
        return None
    fv, vis = pkg("flyvis"), pkg("visual")
    m = fv.parse_flyvis(vdir)
    with open(os.path.join(vdir, "flyvis_map.json")) as f:
        fmap = json.load(f)
    z = np.load(bpath, allow_pickle=False)
    b = pkg("connectome").load_compact(bpath)
    p = dict(b["params"])
    cns = {"N": b["N"], "indptr": b["indptr"], "indices": b["indices"], "counts": z["counts"],
           "sign": z["pre_sign"]}
    steps, pre = 1200, 400
    dirs = [np.asarray(fmap["eyes"][sd]["dirs"], np.float64) for sd in ("L", "R")]
    c = np.array([math.cos(0.17) * math.cos(0.7), math.cos(0.17) * math.sin(0.7), math.sin(0.17)])

    def loom(eye, t_epoch):
        t = (round(t_epoch / 0.02) * 40 - pre) / 2000.0
        if t < 0:
            return np.full(721, 0.5, np.float32)
        rad = (5 + 70 * max(0.0, min(1.0, t / 0.4)) ** 2) * math.pi / 180
        ang = np.arccos(np.minimum(1.0, dirs[eye] @ c))
        return np.where(ang < rad, 0.05, 0.5).astype(np.float32)
    t0 = time.perf_counter()
    ch = vis.VisualChainHip(cns, m, fmap, {**p, "soft": False, "surrogateBeta": 5.0, "sensoryMask": z["sensory"]})
    ch.net.set_input("outScale", np.ones(b["N"], np.float32))
    drive = np.zeros(b["N"], np.float32)
    drive[b["orn"]] = 6.0
    ch.net.set_input("drive", drive)
    _, first = np.unique(b["orn"], return_index=True)         # setDrive's set: first occurrence keeps its place
    ch.net.set_driven(b["orn"][np.sort(first)])
    target = np.concatenate([b[k] for k in ("type:MN9", "type:GNG232", "type:DNge080")]).astype(np.int64)
    dL = np.zeros(b["N"], np.float32)
    dL[target] = 1.0
    ch.net.set_loss(dL, pre)
    ch.settle()
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    L = ch.forward(steps, seed=1, lum=loom)
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = ch.backward(None, lum=True)
    back = time.perf_counter() - t0
    st = ch.stats()
    out = {"workload": "the visual chain at full scale: both eyes of the flyvis model (45,669 nodes each), 62,157 "
                       "eye-to-CNS pairs, the calibrated 165,122-neuron brain as the CNS (hard threshold), 1,200 "
                       "steps (400 of grey, then a looming disc), loss on three cell types' spikes after onset",
           "build_and_settle_s": build_s, "forward_ms": fwd * 1e3, "backward_ms": back * 1e3,
           "ms_per_step_forward": fwd * 1e3 / steps, "ms_per_step_backward": back * 1e3 / steps,
           "loss": L, "live": st["live"], "below_deadband": st["belowDeadband"], "mean_rate_hz": st["meanRate"],
           "grad_gain": g["gain"], "nonzero_fv_weight_grads": int(np.count_nonzero(g["fv"]["weight"])),
           "clamped": [int(x) for x in g["clamped"]]}
    ch.destroy()
    return out


def bench_cord_ensemble(args, brain_per_run: bool = True):
    """The cord lane ensemble (cordens.bend): 64 members x 6 perturbations of
    the reference's ventral-cord brain, 384 runs of 1,000 steps with the
    proprioceptive loop closed on the device; every run's meanHz is checked
    against the reference's published readings.  The reference quotes 29.5 s
    of wall time for the 384 runs on 12 CPU cores (docs/46-cord-ir.md:169-170).
    Skipped when the fixture is absent."""
    import numpy as np
    cdir = os.path.join(ROOT, "tests", "golden", "_cord")
    path, pub = os.path.join(cdir, "cord_brains.npz"), os.path.join(cdir, "cord_search.json")
    if not (os.path.exists(path) and os.path.exists(pub)):
        return None
    z = np.load(path, allow_pickle=False)
    if "ens:poolId" not in z.files:
        return None
    with open(pub) as f:
        pubd = json.load(f)
    want = {(r["member"], ci): rd["meanHz"] for r in pubd["members"] for ci, rd in enumerate(r["reads"])}
    ens = pkg("cordens")
    e = ens.CordEnsemble(z, json.loads(str(z["params"])))
    e.run_lanes(ens.run_order()[:8], steps=50)      # warm-up: first launches of every kernel
    t0 = time.perf_counter()
    got = e.run_lanes()
    wall = time.perf_counter() - t0
    match = sum(ens.mean_hz(x) == want[mc] for mc, x in got.items())
    out = {"workload": "the cord lane ensemble: 64 members x 6 perturbations of the reference's ventral-cord "
                       "brain (21,843 cells), 384 runs of 1,000 steps, proprioceptive loop closed on the device; "
                       "the runs as lanes of the lane ensemble engine, 64 to a launch",
           "runs": len(got), "wall_s": wall, "runs_per_s": len(got) / wall,
           "steps_per_s": len(got) * 1000 / wall, "meanHz_matches_published": match,
           "reference_wall_s_12_cores": 29.5, "speedup_vs_reference": 29.5 / wall}
    if not brain_per_run:
        e.release()
        return out
    e.run(ens.run_order()[:8])                       # warm-up of the one-brain-per-run path
    t0 = time.perf_counter()
    got1 = e.run()
    out["brain_per_run_wall_s"] = time.perf_counter() - t0
    out["brain_per_run_matches_lanes"] = all(np.array_equal(got1[mc], got[mc]) for mc in got)
    e.release()
    return out


def bench_config_run(args, lif, hiprt, stream, hg, p, drv, rates, g):
    """The configuration's whole run as one call: a fresh brain from rest,
    10,000 steps (1 s simulated), on the automatic engine."""
    b = lif.LIFHip(hg, p, seed=1)
    for i, r in zip(drv.tolist(), rates.tolist()):
        b.setDriveOne(i, r)
    b.set_stream(stream.handle)
    b.advance(0)                          # the driven list is uploaded outside the timed call
    e0, e1 = hiprt.Event(), hiprt.Event()
    n = args.config_steps
    c0 = b.counters()
    e0.record(stream)
    b.run_async(n)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    c1 = b.counters()
    ev = c1["edge_events"] - c0["edge_events"]
    out = {"steps": n, "ms": ms, "ms_per_step": ms / n, "edge_events_per_s": ev / (ms / 1e3),
           "sim_s_per_wall_s": n * p["dt"] / 1e3 / (ms / 1e3),
           "spikes_per_step": (c1["spikes"] - c0["spikes"]) / n,
           "launches": int(b.stat("persistent_launches")), "kernel_ms": b.stat("persistent_kernel_ms"),
           "fallback_steps": int(b.stat("fallback_steps")),
           "engine": "persistent" if b.stat("persistent_launches") > 0 else "per-step",
           "workload": "the same brain from rest: all 10,000 steps of the configuration in one call"}
    b.destroy()
    for h in (e0, e1):
        h.destroy()
    return out


def bench_coba(args, lif, hiprt, stream, hg, P, drv, rates, g):
    """The conductance-based variant of the full-scale forward (E_exc 0 mV,
    E_inh -80 mV), timed in steady state on the automatic engine."""
    p = P.config_params(coba=True)
    b = lif.LIFHip(hg, p, seed=1)
    for i, r in zip(drv.tolist(), rates.tolist()):
        b.setDriveOne(i, r)
    b.set_stream(stream.handle)
    b.advance(args.prerun)
    e0, e1 = hiprt.Event(), hiprt.Event()
    n = args.coba_steps
    c0 = b.counters()
    l0 = b.stat("persistent_launches")
    e0.record(stream)
    b.run_async(n)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    c1 = b.counters()
    ev = c1["edge_events"] - c0["edge_events"]
    out = {"steps": n, "ms_per_step": ms / n, "edge_events_per_s": ev / (ms / 1e3),
           "sim_s_per_wall_s": n * p["dt"] / 1e3 / (ms / 1e3), "spikes_per_step": (c1["spikes"] - c0["spikes"]) / n,
           "engine": "persistent" if b.stat("persistent_launches") > l0 else "per-step",
           "workload": f"the full-scale forward, COBA (E_exc 0 mV, E_inh -80 mV), {n} steps after {args.prerun} "
                       "untimed"}
    b.destroy()
    for h in (e0, e1):
        h.destroy()
    return out


def bench_engines(args, lif, hiprt, stream, hg, p, drv, rates):
    """Every engine on the same steady-state workload: a long call (whole slot
    cycles, so the graph engine replays graphs captured beforehand) and a
    call of the timed length.  Graph capture and instantiation happen in
    untimed calls before each measurement."""
    out = {}
    long_n = 76 * 40                      # 3,040 steps: whole 76-step graph replays, same head before and after
    for eng in ("persistent", "graph", "launch"):
        b2 = lif.LIFHip(hg, p, seed=1, engine=eng)
        for i, r in zip(drv.tolist(), rates.tolist()):
            b2.setDriveOne(i, r)
        b2.set_stream(stream.handle)
        b2.advance(args.prerun, engine="auto")
        a0, a1 = hiprt.Event(), hiprt.Event()
        b2.advance(long_n)                  # captures every graph the timed call replays
        a0.record(stream)
        b2.run_async(long_n)
        a1.record(stream)
        a1.synchronize()
        res = {"ms_per_step": a0.elapsed_ms(a1) / long_n, "steps": long_n}
        k = args.steps
        for _ in range(b2.nslots):          # every start head of a k-step call, captured
            b2.advance(k)
        a0.record(stream)
        b2.run_async(k)
        a1.record(stream)
        a1.synchronize()
        res[f"ms_per_step_{k}_step_call"] = a0.elapsed_ms(a1) / k
        out[eng] = res
        b2.destroy()
        for h in (a0, a1):
            h.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--engine", default="auto", choices=["auto", "launch", "graph", "persistent"])
    ap.add_argument("--coba", action="store_true")
    ap.add_argument("--prerun", type=int, default=2000,
                    help="untimed steps from rest to steady state before the warm-up (also the CPU baseline's)")
    ap.add_argument("--config-steps", type=int, default=10000, help="steps of the whole-run leg")
    ap.add_argument("--coba-steps", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=3000, help="steps per CPU baseline sample (median of 3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-engines", action="store_true", help="skip the per-step engine comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference kernel for "
                         "scaling HBM counters), then exit")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-cpu-steps", type=int, default=1000)
    ap.add_argument("--adjoint-batch-cpu-steps", type=int, default=400)
    ap.add_argument("--only", choices=["headline", "flyvis", "visual", "cordens", "lanes", "brain", "cord",
                                       "adjoint", "adjoint32", "batch"],
                    help="run one leg alone and print its JSON (headline: the steady-state forward only; "
                         "adjoint: the full-scale forward + adjoint; adjoint32: its 32-instance batch; "
                         "batch: the 64-member ensemble; cord: the ventral-cord brain)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not measured "
              "(multi-GPU sharding is design only).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 3}))
        return
    if args.only == "flyvis":
        print(json.dumps({"flyvis": bench_flyvis(args)}), flush=True)
        return
    if args.only == "visual":
        print(json.dumps({"visual_chain": bench_visual(args)}), flush=True)
        return
    if args.only == "cordens":
        print(json.dumps({"cord_ensemble": bench_cord_ensemble(args)}), flush=True)
        return
    if args.only == "lanes":
        print(json.dumps({"cord_ensemble_lanes": bench_cord_ensemble(args, brain_per_run=False)}), flush=True)
        return
    if args.only in ("brain", "cord"):
        lif, hiprt = pkg("lif"), pkg("hiprt")
        st = hiprt.Stream()
        if args.only == "brain":
            print(json.dumps({"real_brain": bench_brain(args, lif, hiprt, st)}), flush=True)
        else:
            print(json.dumps({"real_cord": bench_cord(args, lif, hiprt, st)}), flush=True)
        st.destroy()
        return
    synth, P, lif, hiprt, batch, diff = (pkg("synth"), pkg("params"), pkg("lif"), pkg("hiprt"), pkg("batch"),
                                         pkg("diff"))
    if args.only == "batch":
        st = hiprt.Stream()
        print(json.dumps({"batch64": bench_batch64(args, synth, P, lif, batch, hiprt, st)}), flush=True)
        st.destroy()
        return
    if args.only == "adjoint32":
        print(json.dumps({"fwd_adjoint_batch32": bench_adjoint_batch(args, synth, P, diff, hiprt)}), flush=True)
        return
    if args.only == "adjoint":
        st = hiprt.Stream()
        print(json.dumps({"fwd_adjoint": bench_adjoint(args, synth, P, diff, hiprt, st)}), flush=True)
        st.destroy()
        return
    t_build = time.perf_counter()
    g = synth.config2_graph()
    drv = synth.driven_set(g.N, 0.01, 4)
    rates = synth.drive_rates(drv.size, 100.0, 200.0, 5)
    p = P.config_params(coba=args.coba)
    hg = lif.HipGraph.from_synth(g)
    brain = lif.LIFHip(hg, p, seed=1, engine=args.engine)
    for i, r in zip(drv.tolist(), rates.tolist()):
        brain.setDriveOne(i, r)
    print(f"# graph N={g.N} E={g.E} synapses={int(g.counts.sum())} built in {time.perf_counter() - t_build:.1f}s",
          file=sys.stderr, flush=True)

    stream = hiprt.Stream()
    e0, e1 = hiprt.Event(), hiprt.Event()
    brain.set_stream(stream.handle)
    # untimed: the network from rest to its steady state (the first arrivals
    # land D = 18 steps in; the rates settle within a few hundred steps)
    c_pre = brain.counters()
    w_pre = time.perf_counter()
    brain.advance(args.prerun, engine="auto")
    prerun = {"steps": args.prerun, "wall_ms": (time.perf_counter() - w_pre) * 1e3,
              "spikes_per_step": (brain.counters()["spikes"] - c_pre["spikes"]) / max(1, args.prerun),
              "engine": "persistent" if brain.stat("persistent_launches") > 0 else "per-step"}
    brain.advance(args.warmup)
    c0 = brain.counters()
    k0 = brain.stat("persistent_kernel_ms")
    l0 = brain.stat("persistent_launches")
    f0 = brain.stat("fallback_steps")
    hiprt.device_synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    brain.run_async(args.steps)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_ms(e1)
    brain.flush()
    c1 = brain.counters()
    kms = brain.stat("persistent_kernel_ms") - k0
    phases = {f"{blk}_{ph}": brain.stat(f"phase_us_{blk}_{ph}")
              for blk in ("drv", "reg", "spare")
              for ph in ("gather", "sim", "push", "lists", "barrier", "prologue", "share", "pushwait")}
    phases.update({f"drvstep_{ph}": brain.stat(f"phase_us_drvstep_{ph}")
                   for ph in ("inputs", "scan", "draws", "decide", "membrane", "sync")})
    phases.update({f"share_{ph}": brain.stat(f"phase_us_share_{ph}")
                   for ph in ("load", "edges", "publish", "max")})
    overflow_flags = int(brain.stat("persistent_overflow_flags"))
    launches = int(brain.stat("persistent_launches") - l0)
    fallback = int(brain.stat("fallback_steps") - f0)
    events = c1["edge_events"] - c0["edge_events"]
    syn = c1["synapse_events"] - c0["synapse_events"]
    spikes = c1["spikes"] - c0["spikes"]
    secs = ms / 1000.0
    value = events / secs
    bytes_ = algorithmic_bytes(g.N, events, spikes, args.steps)
    engine = "persistent" if launches else ("launch" if args.engine == "auto" else args.engine)
    if launches and not fallback:
        # the lookahead kernel is the only kernel of the timed region: its own
        # HIP-event duration, averaged over launches
        kern_s = kms / 1000.0
        kernel_name = "k_persist (lookahead engine)"
    else:
        kern_s = secs
        kernel_name = "per-step kernel sequence (k_deliver_driven, k_long, k_membrane, k_compact)"
    achieved = bytes_ / kern_s / 1e9
    traffic = None
    if launches == 1 and not fallback:
        # measured HBM bytes of this kernel for a launch of the same length
        # (copy-calibrated counter passes)
        traffic = traffic_per_launch(args.steps)
    out = {
        "metric": "synaptic events/s, full 165k LIF: plain vs batch-64 vs fwd+adjoint",
        "value": value,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "full-scale forward: 165,122 neurons, ~8.0M edges, ~104M synapses, "
                        f"{'COBA' if args.coba else 'CUBA'}, dt 0.1 ms, 1% driven at U(100,200) Hz; timed in steady "
                        f"state after an untimed {args.prerun}-step run from rest",
            "neurons": g.N, "edges": g.E, "synapses": int(g.counts.sum()), "engine": engine,
        },
        "steady_state_prerun": prerun,
        "sim_s_per_wall_s": args.steps * p["dt"] / 1000.0 / secs,
        "synapse_events_per_s": syn / secs,
        "spikes_per_step": spikes / args.steps,
        "edge_events_per_step": events / args.steps,
        "mean_rate_hz": spikes / args.steps / g.N / (p["dt"] / 1000.0),
        "wall_s_host_clock": wall,
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": "profiles/k_persist_hbm_counters.md: rocprofv3 FETCH_SIZE + WRITE_SIZE of the timed "
                              "k_persist dispatch, each scaled by the counter's reading for a 256 MiB copy"
                              if traffic else None,
            "kernel": kernel_name,
            "bytes_per_launch": bytes_ / max(1, launches),
            "launch_ms": kms / max(1, launches) if launches else secs * 1000.0,
            "launches": launches,
            "note": "algorithmic bytes = 32 N per step (v, gE, gI, refr read+written) + 8 per edge event + "
                    "4 per spike; the lookahead kernel keeps neuron state in registers across an epoch, "
                    "so its HBM traffic is far below the algorithmic bytes",
        },
    }
    if args.only == "headline":
        out["lookahead_phases_us_total"] = phases
        brain.destroy()
        hg.release()
        for h in (e0, e1, stream):
            h.destroy()
        print(json.dumps(out), flush=True)
        return
    out["config_run"] = bench_config_run(args, lif, hiprt, stream, hg, p, drv, rates, g)
    out["coba"] = bench_coba(args, lif, hiprt, stream, hg, P, drv, rates, g)
    out["lookahead_phases_us_total"] = phases
    out["lookahead_fallback_steps"] = fallback
    out["lookahead_overflow_flags"] = overflow_flags
    if not args.no_engines:
        out["engines"] = bench_engines(args, lif, hiprt, stream, hg, p, drv, rates)
    brain.destroy()
    cord = bench_cord(args, lif, hiprt, stream)
    if cord:
        out["real_cord"] = cord
    rb = bench_brain(args, lif, hiprt, stream)
    if rb:
        out["real_brain"] = rb
    vis = bench_flyvis(args)
    if vis:
        out["flyvis"] = vis
    chain = bench_visual(args)
    if chain:
        out["visual_chain"] = chain
    cens = bench_cord_ensemble(args)
    if cens:
        out["cord_ensemble"] = cens
    if not args.no_batch:
        out["batch64"] = bench_batch64(args, synth, P, lif, batch, hiprt, stream)
    if not args.no_adjoint:
        out["fwd_adjoint"] = bench_adjoint(args, synth, P, diff, hiprt, stream)
        out["fwd_adjoint_batch32"] = bench_adjoint_batch(args, synth, P, diff, hiprt)
    if not args.no_cpu:
        cb = cpu_baseline(g, p, drv, rates, args.cpu_steps, args.prerun)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_over_cpu_sim_rate"] = out["sim_s_per_wall_s"] / cb["sim_s_per_wall_s"]
            if out.get("config_run"):
                out["config_run"]["gpu_over_cpu_sim_rate"] = (out["config_run"]["sim_s_per_wall_s"]
                                                              / cb["sim_s_per_wall_s"])
    hg.release()
    for h in (e0, e1, stream):
        h.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
