"""GPU parity of the ensemble engine: every member of a batch is bit-identical
to a single brain (the CPU oracle) with its own seed, per-type-gain weights,
silenced neurons (thr = 1e6) and run length."""
import importlib

import numpy as np
import pytest

from oracle import OracleBrain

pytestmark = pytest.mark.gpu

lif = importlib.import_module("fly-brain_amd.lif")
batch = importlib.import_module("fly-brain_amd.batch")
synth = importlib.import_module("fly-brain_amd.synth")
P = importlib.import_module("fly-brain_amd.params")


def member_weights(g, type_of, gains_b):
    """writeGraph with outScale[j] = gain of j's type: W = f32(count * gain)."""
    pre = np.repeat(np.arange(g.N), np.diff(g.indptr.astype(np.int64)))
    return (g.counts.astype(np.float64) * gains_b[type_of[pre]].astype(np.float64)).astype(np.float32)


def run_case(g, sign, p, seeds, type_of, gains, drive_ix, rate, silenced, steps, bias=None, engine="auto"):
    B = len(seeds)
    hg = lif.HipGraph(g.N, g.indptr, g.indices, g.weights, sign, g.counts)
    eb = batch.LIFBatch(hg, p, seeds, type_of, gains)
    eb.set_engine(engine)
    eb.setDrive(drive_ix, rate)
    for b in range(B):
        if len(silenced[b]):
            eb.silence(b, silenced[b])
        if bias is not None:
            eb.setBias(b, bias[0], bias[1])
    eb.set_steps(steps)
    T = int(max(steps))
    nact, ids, masks, fm = eb.run_record(T, cap=T * g.N)
    nf_b, ras_b = eb.member_rasters(nact, ids, masks, fm)
    rngs = eb.rng()
    for b in range(B):
        W = member_weights(g, type_of, gains[b])
        o = OracleBrain(g, p, seed=seeds[b], W=W, sign=sign)
        o.setDrive(drive_ix.tolist(), rate)
        for i in silenced[b]:
            o.setThr(int(i), 1e6)
        if bias is not None:
            o.setBias(bias[0], bias[1])
        nf, ras, _ = o.run(int(steps[b]), raster_cap=int(steps[b]) * g.N)
        assert np.array_equal(nf_b[b, : steps[b]], nf), f"member {b}: per-step counts"
        assert not nf_b[b, steps[b]:].any(), f"member {b}: spikes after its run ended"
        assert np.array_equal(ras_b[b], ras), f"member {b}: raster"
        for k in ("v", "gE", "gI", "refr", "trace", "adapt", "res"):
            assert np.array_equal(eb.array(k, b).view(np.uint32), o.array(k).view(np.uint32)), f"member {b}: {k}"
        assert np.array_equal(eb.array("spikeCount", b), o.spikeCount)
        assert int(rngs[b]) == o.rng, f"member {b}: rng"
    return eb


def _small(seed):
    g = synth.tiny_graph(512, 30, seed=seed, p_count=0.3)
    rng = np.random.default_rng(3)
    B, T = 8, 10
    type_of = rng.integers(0, T, size=g.N).astype(np.int32)
    gains = np.exp(rng.normal(0, 0.3, size=(B, T))).astype(np.float32)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.5).astype(np.float32)
    silenced = [np.sort(rng.choice(g.N, size=int(rng.integers(0, 25)), replace=False)) for _ in range(B)]
    steps = rng.integers(300, 800, size=B)
    return g, B, type_of, gains, sign, silenced, steps


def test_small_batch_members_match_single_brains():
    """Adaptation and depression on: the per-step kernels."""
    g, B, type_of, gains, sign, silenced, steps = _small(41)
    p = P.config_params(adaptInc=0.5, depU=0.1)
    run_case(g, sign, p, [b + 1 for b in range(B)], type_of, gains, np.arange(0, 512, 6), 350.0, silenced, steps,
             bias=(list(range(0, 512, 3)), 6.0))


@pytest.mark.parametrize("engine", ["launch", "persistent"])
def test_small_batch_both_engines_match_single_brains(engine):
    """The lookahead engine and the per-step kernels on the same ensemble,
    with biased neurons and uneven run lengths."""
    g, B, type_of, gains, sign, silenced, steps = _small(43)
    p = P.config_params()
    eb = run_case(g, sign, p, [b + 1 for b in range(B)], type_of, gains, np.arange(0, 512, 6), 350.0, silenced,
                  steps, bias=(list(range(0, 512, 3)), 6.0), engine=engine)
    if engine == "persistent":
        assert eb.stat("lookahead_steps") > 0


def test_lookahead_ensemble_bucket_overflow_falls_back():
    """A saturated ensemble overflows the lookahead buckets; those epochs run
    on the per-step kernels and every member stays exact."""
    g = synth.tiny_graph(2048, 200, seed=31, p_count=0.5)
    B = 8
    type_of = np.zeros(g.N, np.int32)
    gains = np.ones((B, 1), np.float32)
    sign = (np.where(g.exc, 1.0, -1.0) * 1.5).astype(np.float32)
    silenced = [np.zeros(0, np.int64) for _ in range(B)]
    steps = np.full(B, 400)
    eb = run_case(g, sign, P.config_params(), [b + 1 for b in range(B)], type_of, gains, np.arange(0, 2048, 10),
                  400.0, silenced, steps, bias=(list(range(2048)), 6.5), engine="persistent")
    assert eb.stat("fallback_steps") > 0


@pytest.mark.parametrize("engine", ["launch", "persistent"])
def test_batch64_config3_members_match_single_brains(engine):
    """The ensemble workload: the 16,384-neuron graph, B = 64, 500 types with
    lognormal(0, 0.2) gains, 0-5% silenced per member, uneven run lengths
    (shortened here so 64 oracle runs finish in seconds)."""
    g = synth.config1_graph()
    B = 64
    type_of = synth.type_ids(g.N, 500, 6)
    gains = np.stack([synth.type_gains(b) for b in range(B)])
    silenced = [synth.silenced(b, g.N) for b in range(B)]
    steps = np.array([200 + synth.run_length(b) // 10 for b in range(B)])
    drv = synth.driven_set(g.N, 0.02, 3)
    eb = run_case(g, g.sign, P.config_params(), [b + 1 for b in range(B)], type_of, gains, drv, 150.0, silenced,
                  steps, engine=engine)
    c = eb.counters()
    assert c["edge_events"] > 0 and c["spikes"] > 0


def test_engine_switches_and_drive_changes_keep_members_exact():
    """Runs alternate between the per-step kernels and the lookahead engine
    (the delay line is handed over both ways, at lengths that are not whole
    epochs), and the driven set changes between runs; every member stays
    bit-identical to its single-brain run with the same calls."""
    g = synth.tiny_graph(3000, 40, seed=53, p_count=0.3)
    B = 21
    rng = np.random.default_rng(9)
    type_of = rng.integers(0, 7, size=g.N).astype(np.int32)
    gains = np.exp(rng.normal(0, 0.2, size=(B, 7))).astype(np.float32)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.55).astype(np.float32)
    p = P.config_params()
    hg = lif.HipGraph(g.N, g.indptr, g.indices, g.weights, sign, g.counts)
    eb = batch.LIFBatch(hg, p, [b + 7 for b in range(B)], type_of, gains)
    oracles = []
    for b in range(B):
        o = OracleBrain(g, p, seed=b + 7, W=member_weights(g, type_of, gains[b]), sign=sign)
        oracles.append(o)
    drv1 = np.arange(0, 3000, 13)
    drv2 = np.arange(5, 3000, 29)
    plan = [("drive", drv1, 300.0), ("launch", 55), ("persistent", 200), ("launch", 17), ("persistent", 91),
            ("drive", drv2, 260.0), ("drive", drv1[:40], 0.0), ("persistent", 150), ("launch", 40)]
    nf_all = [[] for _ in range(B)]
    for item in plan:
        if item[0] == "drive":
            eb.setDrive(item[1], item[2])
            for o in oracles:
                o.setDrive(item[1].tolist(), item[2])
            continue
        engine, n = item
        eb.set_engine(engine)
        nact, ids, masks, fm = eb.run_record(n, cap=n * g.N)
        nf_b, ras_b = eb.member_rasters(nact, ids, masks, fm)
        for b in range(B):
            nf, ras, _ = oracles[b].run(n, raster_cap=n * g.N)
            assert np.array_equal(nf_b[b], nf), (engine, n, b)
            assert np.array_equal(ras_b[b], ras), (engine, n, b)
    assert eb.stat("lookahead_steps") > 0
    rngs = eb.rng()
    for b in range(B):
        for k in ("v", "gE", "gI", "refr", "trace"):
            assert np.array_equal(eb.array(k, b).view(np.uint32), oracles[b].array(k).view(np.uint32)), (b, k)
        assert np.array_equal(eb.array("spikeCount", b), oracles[b].spikeCount)
        assert int(rngs[b]) == oracles[b].rng


def test_lane_drive_cannot_be_cleared_under_the_pool_loop():
    """The pool loop writes every lane's drive after each step: switching
    per-lane drive off while a loop is set is refused (no step runs), and it
    is allowed again once the loop is cleared."""
    print("lane drive vs pool loop: argument check only", flush=True)
    g = synth.tiny_graph(256, 8, seed=3)
    hg = lif.HipGraph.from_synth(g)
    bt = batch.LIFBatch(hg, P.config_params(), [1, 2], np.zeros(g.N, np.int32), np.ones((2, 1), np.float32))
    bt.set_lanes(True, np.zeros((2, g.N), np.float32))
    pool = np.full(g.N, 255, np.uint8)
    pool[:6] = np.arange(6)
    bt.set_pool_feedback(pool, 6, np.arange(10, 20, dtype=np.int32), 20.0, 1.0, [2, 3], 10)
    with pytest.raises(RuntimeError, match="pool loop"):
        bt.set_lanes(True, None)
    bt.set_pool_feedback(None, 6, np.zeros(0, np.int32), 0.0, 0.0, [1, 1], 0)
    bt.set_lanes(True, None)
    bt.set_lanes(True, np.ones((2, g.N), np.float32))     # the kept buffer is reused
    bt.destroy()
    hg.release()


def test_every_lookahead_shape_and_the_autotuner_give_the_same_bits():
    """The ensemble lookahead engine cut into regular blocks of 72, 80, 96 and
    128 neurons (fb_batch_set_shape), and the autotuner's own choice: every member
    equals its single-brain oracle run for each, and the autotuner reports
    the shape it chose and what it measured for every candidate."""
    g = synth.tiny_graph(4096, 40, seed=47, p_count=0.3)
    rng = np.random.default_rng(9)
    B, T = 64, 40
    type_of = rng.integers(0, T, size=g.N).astype(np.int32)
    gains = np.exp(rng.normal(0, 0.2, size=(B, T))).astype(np.float32)
    sign = (np.where(g.exc, 1.0, -1.0) * 0.5).astype(np.float32)
    silenced = [np.sort(rng.choice(g.N, size=int(rng.integers(0, 150)), replace=False)) for _ in range(B)]
    steps = rng.integers(400, 700, size=B)
    p = P.config_params()
    for shape in (72, 80, 96, 128, 0):
        print(f"ensemble lookahead shape {shape or 'autotuned'}", flush=True)
        hg = lif.HipGraph(g.N, g.indptr, g.indices, g.weights, sign, g.counts)
        eb = batch.LIFBatch(hg, p, [b + 1 for b in range(B)], type_of, gains)
        eb.set_engine("persistent")
        eb.set_shape(shape)
        eb.setDrive(np.arange(0, g.N, 23), 150.0)
        for b in range(B):
            if len(silenced[b]):
                eb.silence(b, silenced[b])
        eb.set_steps(steps)
        eb.run(int(steps.max()))
        if shape:
            assert eb.stat("lookahead_shape") == shape
        else:
            # a shape whose trial epochs overflowed its buckets has no measurement
            chosen = int(eb.stat("lookahead_shape"))
            assert chosen in (72, 80, 96, 128)
            assert eb.stat(f"autotune_us_per_epoch_{chosen}") > 0
        for b in range(0, B, 9):
            W = member_weights(g, type_of, gains[b])
            o = OracleBrain(g, p, seed=b + 1, W=W, sign=sign)
            o.setDrive(np.arange(0, g.N, 23).tolist(), 150.0)
            for i in silenced[b]:
                o.setThr(int(i), 1e6)
            o.run(int(steps[b]), raster_cap=int(steps[b]) * g.N)
            for k in ("v", "gE", "gI", "refr"):
                assert np.array_equal(eb.array(k, b).view(np.uint32), o.array(k).view(np.uint32)), (shape, b, k)
            assert np.array_equal(eb.array("spikeCount", b), o.spikeCount), (shape, b)
        if shape == 80:
            assert eb.stat("fallback_steps") == 0       # larger blocks may overflow their buckets: exact anyway
        eb.destroy()
        hg.release()


def test_batch64_config3_full_length_members():
    """The ensemble workload at its own run lengths (2,000-10,000 steps,
    U{2000..10000} per member) on the automatic engine (the lookahead kernel
    with its autotuned shape): the members with the shortest and the longest
    runs, and two more, equal their single-brain oracle runs -- state, spike
    counts and stream -- after the whole run."""
    g = synth.config1_graph()
    B = 64
    type_of = synth.type_ids(g.N, 500, 6)
    gains = np.stack([synth.type_gains(b) for b in range(B)])
    silenced = [synth.silenced(b, g.N) for b in range(B)]
    steps = np.array([synth.run_length(b) for b in range(B)])
    drv = synth.driven_set(g.N, 0.02, 3)
    p = P.config_params()
    print(f"C3 at full length: runs {steps.min()}..{steps.max()} steps", flush=True)
    hg = lif.HipGraph.from_synth(g)
    eb = batch.LIFBatch(hg, p, [b + 1 for b in range(B)], type_of, gains)
    eb.setDrive(drv, 150.0)
    for b in range(B):
        if len(silenced[b]):
            eb.silence(b, silenced[b])
    eb.set_steps(steps)
    eb.run(int(steps.max()))
    assert eb.stat("lookahead_steps") > 0
    rngs = eb.rng()
    pick = sorted({int(np.argmin(steps)), int(np.argmax(steps)), 1, 33})
    for b in pick:
        o = OracleBrain(g, p, seed=b + 1, W=member_weights(g, type_of, gains[b]), sign=g.sign)
        o.setDrive(drv.tolist(), 150.0)
        for i in silenced[b]:
            o.setThr(int(i), 1e6)
        o.run(int(steps[b]), raster_cap=1)
        for k in ("v", "gE", "gI", "refr", "trace"):
            assert np.array_equal(eb.array(k, b).view(np.uint32), o.array(k).view(np.uint32)), (b, k)
        assert np.array_equal(eb.array("spikeCount", b), o.spikeCount), b
        assert int(rngs[b]) == o.rng, b
    eb.destroy()
    hg.release()
