"""Writes tests/golden/_cord/cord_brains.npz: the reference's ventral-cord
brain (scripts/cord_kernel_ref.mjs) built from its data files by
fly-brain_amd/connectome.py, for the GPU test that replays the reference's
published spike totals (docs/46-cord-ir.md:93-96), with the lane ensemble's
tables (scripts/cord_ens_tab.mjs); and a copy of the published ensemble
readings (public/data/cord_search.json).

Run where the reference's data exist (`__graft_entry__.build()` does it):
    python tests/golden/make_cord_fixture.py [/root/reference/public/data]
The output is derived data and stays out of git (tests/golden/_cord/ is
ignored); it travels with the working tree to the GPU machine.
"""
import importlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
C = importlib.import_module("fly-brain_amd.connectome")

ABLATIONS = ("hemilineage:13A", "hemilineage:13B", "hemilineage:19B")


def main(data_dir="/root/reference/public/data"):
    if not os.path.exists(os.path.join(data_dir, "cord_ir.bin")):
        sys.exit(f"no cord data under {data_dir}")
    b = C.cord_brain(data_dir)
    cord = C.load_cord(data_dir)
    out = os.path.join(HERE, "_cord")
    os.makedirs(out, exist_ok=True)
    arrays = {"indptr": b["indptr"], "indices": b["indices"], "W": b["W"], "sign": b["sign"],
              "driven": b["driven"].astype(np.int32), "rates": b["rates"].astype(np.float32)}
    for name in ABLATIONS:
        arrays["silenced:" + name] = cord["sets"][name].astype(np.int32)
    # the lane ensemble's per-cell and per-edge tables (scripts/cord_ens_tab.mjs):
    # leg pool of each cell (6 = none), proprioceptors, hemilineage bits
    # (1 = 13A, 2 = 13B, 4 = 19B), role, and the commissural edges
    pool = np.full(cord["N"], 6, np.uint8)
    for l, name in enumerate(cord["legOrder"]):
        pool[cord["sets"]["pool:" + name]] = l
    arrays["ens:poolId"] = pool
    arrays["ens:proprio"] = cord["sets"]["sensory:proprioceptive"].astype(np.int32)
    arrays["ens:hl"] = cord["hl"].astype(np.uint8)
    arrays["ens:role"] = cord["role"].astype(np.uint8)
    arrays["ens:comm"] = (cord["edgeClass"] & 1).astype(np.uint8)
    # the published ensemble readings (public/data/cord_search.json), the
    # known answer of the ensemble test
    src = os.path.join(data_dir, "cord_search.json")
    if os.path.exists(src):
        shutil.copyfile(src, os.path.join(out, "cord_search.json"))
    params = {k: v for k, v in b["params"].items() if isinstance(v, (int, float, bool))}
    np.savez(os.path.join(out, "cord_brains.npz"), params=np.array(json.dumps(params)), **arrays)
    print("wrote", os.path.join(out, "cord_brains.npz"))


if __name__ == "__main__":
    main(*sys.argv[1:])
