"""Writes tests/golden/lif_golden.json from the reference's own lif.c.

Run in a container that has /root/reference: `make -C oracle ref` first, then
`python tests/golden/make_golden.py`.  Each case of tests/lifcases.py is played
on oracle/_ref (lif.c compiled natively, -O2 -ffp-contract=off) and its
per-step spike counts, raster and final state are recorded as hashes.  The
fixture is data only: seeds, digests and counts.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import lifcases  # noqa: E402
from oracle import OracleBrain, available  # noqa: E402


def main():
    if not available("ref"):
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where /root/reference exists")
    out = {}
    for name in lifcases.CASES:
        case, g, W, s, p = lifcases.build_inputs(name)
        b = OracleBrain(g, p, seed=case["seed"], kind="ref", W=W, sign=s)
        nf, ras = lifcases.play(b, case["ops"], g.N)
        rec = lifcases.summary(b, nf, ras)
        rec["graph_digest"] = g.digest()
        rec["W_sha"] = lifcases.sha(W)
        rec["sign_sha"] = lifcases.sha(s)
        rec["steps"] = int(nf.size)
        rec["nfired_head"] = nf[:64].tolist()
        out[name] = rec
        print(f"{name:20s} steps {nf.size:5d} spikes {rec['total']:7d}")
    with open(os.path.join(HERE, "lif_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
