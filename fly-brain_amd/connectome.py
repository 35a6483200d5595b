"""The real-connectome build path: the reference's preprocessed data files and
the construction of a calibrated brain from them.

Restates, in numpy, the host-side steps the reference runs before its kernel
(the file formats and the order of every f64 operation are theirs):
- ``load_connectome`` reads meta.json, neurons.bin, graph_w<minWeight>.bin,
  neuron_size.bin and ntsign.bin (scripts/lib_node.mjs:3-14);
- ``region_size_ref`` / ``brain_scales`` give the per-target input scale from
  neuron volume against its region's median, and the sensory mask
  (src/ratenet.js:23-31, src/brainmodel.js:8-19);
- ``modulatory_sign`` silences octopaminergic fast synapses when
  neuromodulation is on (src/brainmodel.js:33-41, src/sim/neuromod.js:51);
- ``write_graph`` bakes weights and signs as writeGraph does (f64 products,
  one f32 rounding; src/lifwasm.js:14-34);
- ``load_cord`` / ``cord_brain`` build the 21,843-cell ventral-cord brain of
  scripts/cord_kernel_ref.mjs:28-105 from cord_ir.bin, with its drive and its
  hemilineage ablations.

The data directory is the reference's ``public/data``; nothing here is
needed by the engine itself, which takes the baked arrays.
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

from .params import DEFAULTS

# src/brainmodel.js:7
BRAIN_DEFAULTS = {"laminaBias": 9.0, "kcThreshold": 0.0, "wSyn": 0.3, "sizeAlpha": 0.5, "minSyn": 5,
                  "adaptInc": 0.0, "depU": 0.0, "maxSizeScale": 20.0, "boostCap": 1.0}
# src/lif.js:5
EXC_SIGN = np.array([0, 1, -1, -1, 1, 1, 1, -1], dtype=np.float64)


def calibrated_params(data_dir: str) -> dict:
    """``{...DEFAULTS, ...BRAIN_DEFAULTS, ...brain_params.json}`` (scripts/cord_kernel_ref.mjs:47-50)."""
    with open(os.path.join(data_dir, "brain_params.json")) as f:
        bp = json.load(f)
    p = dict(DEFAULTS)
    p.update(BRAIN_DEFAULTS)
    p.update({k: v for k, v in bp.items() if not k.startswith("_")})
    return p


def load_connectome(data_dir: str) -> dict:
    """The preprocessed connectome (scripts/lib_node.mjs:3-14) plus the per-neuron
    volume and sign files."""
    with open(os.path.join(data_dir, "meta.json")) as f:
        meta = json.load(f)
    N = int(meta["N"])
    nb = np.fromfile(os.path.join(data_dir, "neurons.bin"), dtype=np.uint8)
    off = 8 + N * 8 + N * 12            # header, body ids, soma xyz
    indeg = nb[off:off + 4 * N].view(np.uint32); off += 4 * N
    outdeg = nb[off:off + 4 * N].view(np.uint32); off += 4 * N
    cls = nb[off:off + 2 * N].view(np.uint16); off += 2 * N
    nt = nb[off:off + N].copy(); off += N
    sc = nb[off:off + N].copy(); off += N
    side = nb[off:off + N].copy()
    g = np.fromfile(os.path.join(data_dir, f"graph_w{meta['minWeight']}.bin"), dtype=np.uint8)
    E = int(g[4:8].view(np.uint32)[0])
    o = 8
    indptr = g[o:o + 4 * (N + 1)].view(np.uint32); o += 4 * (N + 1)
    indices = g[o:o + 4 * E].view(np.uint32); o += 4 * E
    weights = g[o:o + 2 * E].view(np.uint16)
    size = np.fromfile(os.path.join(data_dir, "neuron_size.bin"), dtype=np.float32)
    ntsign = np.fromfile(os.path.join(data_dir, "ntsign.bin"), dtype=np.float32)
    return {"meta": meta, "N": N, "E": E, "indeg": indeg, "outdeg": outdeg, "cls": cls, "nt": nt, "sc": sc,
            "side": side, "indptr": indptr, "indices": indices, "weights": weights, "size": size, "ntsign": ntsign}


def region_of(name: str) -> int:
    """src/ratenet.js:23: 0 optic lobe, 2 ventral cord and its tracts, 1 central brain."""
    if re.match(r"^(ol_|visual_projection|visual_centrifugal)", name):
        return 0
    if re.match(r"^(vnc_|descending|ascending|sensory_ascending|sensory_descending|efferent_)", name):
        return 2
    return 1


def region_size_ref(superclasses, sc, size):
    """Per neuron, the median positive volume of its region (src/ratenet.js:24-31):
    the sorted bucket's element at index len >> 1, or 1 for an empty bucket."""
    reg_of_sc = np.array([region_of(n) for n in superclasses], dtype=np.uint8)
    region = reg_of_sc[np.asarray(sc, dtype=np.int64)]
    med = []
    for r in range(3):
        b = np.sort(size[(region == r) & (size > 0)].astype(np.float64))
        med.append(float(b[b.size >> 1]) if b.size else 1.0)
    ref = np.array(med, dtype=np.float32)[region]
    return ref, med, region


def brain_scales(conn: dict, size, opts: dict):
    """(inScale f32[N], sensoryMask u8[N], region u8[N]) as src/brainmodel.js:8-19
    computes them: s = clamp(size / ref, 1/maxSizeScale, maxSizeScale) (1 when
    the volume is unknown), inScale = min(boostCap, s^-sizeAlpha) x vncGain in
    the cord region."""
    o = dict(BRAIN_DEFAULTS)
    o.update(opts)
    names = conn["meta"]["superclasses"]
    ref, _, region = region_size_ref(names, conn["sc"], size)
    size64 = np.asarray(size, dtype=np.float32).astype(np.float64)
    mx = float(o["maxSizeScale"])
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.minimum(mx, np.maximum(1.0 / mx, size64 / ref.astype(np.float64)))
    s = np.where(size64 > 0, s, 1.0)
    v = np.minimum(float(o["boostCap"]), np.power(s, -float(o["sizeAlpha"])))
    vnc = o.get("vncGain")
    if vnc:
        v = v * np.where(region == 2, float(vnc), 1.0)
    in_scale = v.astype(np.float32)
    sensory = np.array([1 if re.search("sensory", names[k]) else 0 for k in range(len(names))], dtype=np.uint8)
    return in_scale, sensory[np.asarray(conn["sc"], dtype=np.int64)], region


def modulatory_sign(conn: dict, pre_sign, opts: dict):
    """src/brainmodel.js:33-41: with neuromodulation on (and no 'synFast' mode),
    the octopaminergic neurons' fast synapses leave the graph (sign 0)."""
    if not opts.get("neuromod"):
        return pre_sign
    base = np.asarray(pre_sign, dtype=np.float32).copy() if pre_sign is not None else \
        EXC_SIGN[np.asarray(conn["nt"], dtype=np.int64)].astype(np.float32)
    if opts.get("oaMode") == "synFast":
        return base
    types = conn["meta"]["types"]
    oa = np.array([bool(re.match(r"^OA-", t or "")) for t in types])
    base[oa] = 0.0
    return base


def write_graph(N: int, indptr, indices, counts, pre_sign, params: dict, in_scale=None, sensory_mask=None,
                out_scale=None, edge_gain=None):
    """(W f32[E], sign f32[N]) as writeGraph bakes them (src/lifwasm.js:14-34):
    ``W[k] = f32(c * edgeGain[k] * inScale[q] * ig)`` with ``ig = (s < 0 ? inhGain : 1)
    * outScale[j]`` for an edge j -> q whose count passes minSyn and whose target
    is not sensory (unless presynaptic inhibition keeps an inhibitory source's
    edges); ``sign[j] = f32(s_j * wSyn)``.  Every product is binary64, rounded
    once."""
    ip = np.asarray(indptr, dtype=np.int64)
    ix = np.asarray(indices, dtype=np.int64)
    c = np.asarray(counts).astype(np.float64)
    E = int(ip[-1])
    s = np.asarray(pre_sign, dtype=np.float64).reshape(N)
    row = np.repeat(np.arange(N), np.diff(ip))
    og = np.ones(N) if out_scale is None else np.asarray(out_scale, dtype=np.float32).astype(np.float64)
    ig = np.where(s < 0, float(params["inhGain"]), 1.0) * og
    eg = np.ones(E) if edge_gain is None else np.asarray(edge_gain, dtype=np.float32).astype(np.float64)
    inS = np.ones(E) if in_scale is None else np.asarray(in_scale, dtype=np.float32).astype(np.float64)[ix]
    sens = np.zeros(E, dtype=bool) if sensory_mask is None else np.asarray(sensory_mask).astype(bool)[ix]
    keep = ~sens | ((float(params.get("preInh") or 0.0) > 0) & (s[row] < 0))
    on = (c >= float(params["minSyn"])) & keep
    W = np.where(on, c * eg * inS * ig[row], 0.0).astype(np.float32)
    sign = (s * float(params["wSyn"])).astype(np.float32)
    return W, sign


def load_cord(data_dir: str) -> dict:
    """cord_ir.bin with its layout in cord_ir.json (scripts/cord_kernel_ref.mjs:28-45):
    the ventral-cord subgraph, its map to full-connectome indices and its named
    cell sets.  Row targets are stored as gaps: index = previous + gap + 1."""
    with open(os.path.join(data_dir, "cord_ir.json")) as f:
        d = json.load(f)
    raw = np.fromfile(os.path.join(data_dir, "cord_ir.bin"), dtype=np.uint8)
    dt = {"origIdx": np.uint32, "indptr": np.uint32, "gaps": np.uint32, "counts": np.uint16, "edgeClass": np.uint8,
          "role": np.uint8, "nt": np.uint8, "side": np.uint8, "hl": np.uint8, "sign": np.float32, "size": np.float32,
          "setPtr": np.uint32, "setIdx": np.uint32}

    def at(name):
        L = d["layout"][name]
        t = np.dtype(dt[name])
        return raw[L["off"]:L["off"] + L["n"] * t.itemsize].view(t)

    N, E = int(d["counts"]["N"]), int(d["counts"]["E"])
    indptr = at("indptr").astype(np.int64)
    gaps = at("gaps").astype(np.int64)
    indices = np.empty(E, dtype=np.int64)
    for k in range(N):
        a, b = indptr[k], indptr[k + 1]
        if b > a:
            indices[a:b] = np.cumsum(gaps[a:b] + 1) - 1
    set_ptr, set_idx = at("setPtr"), at("setIdx")
    sets = {s["name"]: set_idx[set_ptr[i]:set_ptr[i + 1]].astype(np.int64) for i, s in enumerate(d["sets"])}
    return {"desc": d, "N": N, "E": E, "indptr": indptr.astype(np.uint32), "indices": indices.astype(np.uint32),
            "counts": at("counts"), "origIdx": at("origIdx").astype(np.int64), "role": at("role"), "nt": at("nt"),
            "sets": sets}


def cord_brain(data_dir: str, ablate: str | None = None) -> dict:
    """The cord brain of scripts/cord_kernel_ref.mjs:52-105: calibrated
    parameters; input scale, sensory mask and sign from the full-connectome
    build, restricted to the cord's cells; weights baked by writeGraph; 60 Hz
    drive to the descending neurons and 20 Hz to the proprioceptors, in
    ascending index order; an optional named set silenced by threshold."""
    p = calibrated_params(data_dir)
    if p["depU"] != 0:
        raise ValueError("the cord brain assumes depU = 0")
    conn = load_connectome(data_dir)
    cord = load_cord(data_dir)
    in_scale, sensory, _ = brain_scales(conn, conn["size"], p)
    csign = modulatory_sign(conn, conn["ntsign"], p)
    o = cord["origIdx"]
    c_in, c_sens, c_sign = in_scale[o], sensory[o], np.asarray(csign, dtype=np.float32)[o]
    if not np.array_equal(c_sens.astype(bool), cord["role"] == 3):
        raise ValueError("cord roles and the sensory mask disagree")
    W, sign = write_graph(cord["N"], cord["indptr"], cord["indices"], cord["counts"], c_sign, p, c_in, c_sens)
    drive = np.zeros(cord["N"], dtype=np.float32)
    drive[cord["sets"]["role:descending"]] = 60.0
    drive[cord["sets"]["sensory:proprioceptive"]] = 20.0
    driven = np.nonzero(drive > 0)[0]
    silenced = cord["sets"][ablate] if ablate else np.zeros(0, dtype=np.int64)
    return {"N": cord["N"], "indptr": cord["indptr"], "indices": cord["indices"], "W": W, "sign": sign,
            "params": p, "driven": driven, "rates": drive[driven], "silenced": silenced}
