"""Writes tests/golden/_brain/calibrated_brain.npz: the reference's whole
calibrated brain (165,122 neurons, scripts/calib_eval.mjs:43-51) as
fly-brain_amd/connectome.py builds it from the reference's data files, in
``save_compact``'s form, with the sugar assay's drive sets and the neurons of
the types whose rates the reference reports for it.

Run where the reference's data exist (`__graft_entry__.build()` does it):
    python tests/golden/make_brain_fixture.py [/root/reference/public/data]
The output is derived data and stays out of git (tests/golden/_brain/ is
ignored); it travels with the working tree to the GPU machine.
"""
import importlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
C = importlib.import_module("fly-brain_amd.connectome")
FIXTURE = os.path.join(HERE, "_brain", "calibrated_brain.npz")


def main(data_dir="/root/reference/public/data"):
    if not os.path.exists(os.path.join(data_dir, "graph_w3.bin")):
        sys.exit(f"no connectome under {data_dir}")
    conn = C.load_connectome(data_dir)
    with open(os.path.join(data_dir, "bodymap.json")) as f:
        conn["bodymap"] = json.load(f)
    a = C.sugar_assay(conn)
    extra = {"orn": a["orn"].astype(np.int32), "sugar": a["sugar"].astype(np.int32)}
    for t in ("MN9", "GNG232", "DNge080"):
        extra["type:" + t] = C.by_type(conn, t).astype(np.int32)
    os.makedirs(os.path.dirname(FIXTURE), exist_ok=True)
    C.save_compact(FIXTURE, data_dir, extra=extra)
    print("wrote", FIXTURE)


if __name__ == "__main__":
    main(*sys.argv[1:])
