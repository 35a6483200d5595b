"""exact32 (src/lifdiff.js:105-109): the differentiable model's forward pass
in lif.c's binary32 arithmetic is the reference's own lif.c bit for bit --
spike lists, every state array and the stream -- the gate of the reference's
twin audit (scripts/twin_audit.mjs:107-144), here on the oracle."""
import numpy as np
import pytest

import exact32cases as xc
from oracle import available

KIND = "ref" if available("ref") else "oracle"
STATE = (("v", "v"), ("gE", "gE"), ("gI", "gI"), ("adapt", "adapt"), ("res", "res"), ("refr", "refr"))


def _split(off, idx, forced, steps):
    out = []
    for t in range(steps):
        out.append(idx[off[t]:off[t + 1]].tolist())
    return out


@pytest.mark.parametrize("over", [{}, {"depU": 0.0}, {"adaptInc": 0.0, "depU": 0.3},
                                  {"dt": 0.1, "tRef": 2.2}])
def test_exact32_forward_is_lifc_bit_for_bit(over):
    c = xc.Case(over=over)
    steps = 400
    o = c.oracle()
    o.forward(steps, seed=3)
    off, idx, amp, forced = o.spikes()
    b = c.lifc(seed=3, kind=KIND)
    nf, ras, _ = b.run(steps, raster_cap=steps * c.g.N)
    assert np.array_equal(np.diff(off), nf)
    assert np.array_equal(idx, ras)
    assert int(nf.sum()) > 200                           # the run is not silent
    for a, k in STATE:
        assert np.array_equal(o.state(a).view(np.uint32), b.array(k).view(np.uint32)), a
    assert np.array_equal(o.state("spikeCount"), b.spikeCount.astype(np.float32))
    assert o.rng == b.rng


def test_exact32_differs_from_the_binary64_forward():
    """The default forward is LIFDiff's binary64 arithmetic: it is not lif.c
    bit for bit, so the mode is doing the work."""
    c = xc.Case()
    o = c.oracle({"exact32": False})
    o.forward(400, seed=3)
    b = c.lifc(seed=3, kind=KIND)
    b.run(400, raster_cap=400 * c.g.N)
    assert not np.array_equal(o.state("v").view(np.uint32), b.array("v").view(np.uint32))


def test_exact32_refuses_a_threshold_mask():
    """lif.c sums vThresh + adapt + thr only, so exact32 needs the whole
    threshold field in thrOffset (src/lifdiff.js:178-180)."""
    c = xc.Case()
    o = c.oracle()
    m = np.zeros(c.g.N, np.uint8)
    m[5] = 1
    o.set_masks(m, c.bias_mask)
    assert np.isnan(o.forward(10, seed=1))


def test_depression_rounding_is_lifc_not_binary64():
    """lifdiff.js's exact32 line `res[pre] -= dU * res[pre]` rounds once from
    binary64; lif.c rounds the product to binary32 first.  They disagree on a
    sizeable share of updates at depU 0.15 -- why this restatement takes
    lif.c's rounding for the mode whose purpose is lif.c's bits."""
    rng = np.random.default_rng(1)
    d = np.float32(0.15)
    a = rng.uniform(0.01, 1, 200000).astype(np.float32)
    js = (a.astype(np.float64) - np.float64(d) * a.astype(np.float64)).astype(np.float32)
    c = a - (d * a).astype(np.float32)
    assert 0.05 < float(np.mean(js != c)) < 0.2
