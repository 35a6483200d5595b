"""Writes tests/golden/grad_check_subgraph.npz: the induced subgraph of the
real connectome that the reference's gradient check runs on
(scripts/grad_check.mjs:26-44), for K = 800 neurons -- the size its published
table is quoted at (docs/33-differentiable-brain.md:48-62).

The subgraph: the seed is the first neuron of highest out-degree under 400
(grad_check.mjs:27); neurons are taken breadth-first from it in row order
until K are in (:28-35); an edge a -> b is kept when both ends are in, with
its synapse count, and rows keep the connectome's edge order (:37-43).

Run where the reference's data exist:
    python tests/golden/make_grad_check_fixture.py [/root/reference/public/data] [K]
The output is a few hundred KB of derived data and is committed.
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
C = importlib.import_module("fly-brain_amd.connectome")
FIXTURE = os.path.join(HERE, "grad_check_subgraph.npz")


def subgraph(indptr, indices, weights, nt, K):
    N = indptr.size - 1
    outdeg = np.diff(indptr.astype(np.int64))
    seed = 0
    for i in range(N):                      # strictly greater: the first of the highest wins
        if outdeg[i] > outdeg[seed] and outdeg[i] < 400:
            seed = i
    pick, in_set = [seed], {seed: 0}
    h = 0
    while h < len(pick) and len(pick) < K:
        u = pick[h]
        for j in range(int(indptr[u]), int(indptr[u + 1])):
            if len(pick) >= K:
                break
            q = int(indices[j])
            if q not in in_set:
                in_set[q] = len(pick)
                pick.append(q)
        h += 1
    n = len(pick)
    ip = np.zeros(n + 1, np.uint32)
    ix, wt = [], []
    for a, u in enumerate(pick):
        for j in range(int(indptr[u]), int(indptr[u + 1])):
            m = in_set.get(int(indices[j]))
            if m is not None:
                ix.append(m)
                wt.append(int(weights[j]))
        ip[a + 1] = len(ix)
    return (np.array(pick, np.int32), ip, np.array(ix, np.uint32), np.array(wt, np.uint16),
            np.asarray(nt)[np.array(pick)].astype(np.uint8), seed)


def main(data_dir="/root/reference/public/data", K=800):
    if not os.path.exists(os.path.join(data_dir, "graph_w3.bin")):
        sys.exit(f"no connectome under {data_dir}")
    conn = C.load_connectome(data_dir)
    pick, ip, ix, wt, nt, seed = subgraph(conn["indptr"], conn["indices"], conn["weights"], conn["nt"], int(K))
    np.savez_compressed(FIXTURE, pick=pick, indptr=ip, indices=ix, counts=wt, nt=nt, seed=np.int64(seed))
    print(f"wrote {FIXTURE}: {pick.size} neurons, {ix.size} connections, seed neuron {seed}")


if __name__ == "__main__":
    main(*sys.argv[1:])
