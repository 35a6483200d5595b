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
- ``type_gains`` gives the per-neuron outgoing gains of named cell types and
  classes (src/brainmodel.js:58-100);
- ``class_physiology`` and ``neuromod_tone`` give the thresholds and biases
  the calibrated brain starts with: Kenyon-cell thresholds and lamina bias
  (src/brainmodel.js:20-30), then the fed octopamine tone of the
  neuromodulation module (src/sim/neuromod.js:60-160, its constructor and one
  ``modulate()``);
- ``calibrated_brain`` puts these together as scripts/calib_eval.mjs:43-51
  (``makeBrain``) does for the whole 165,122-neuron brain, and
  ``sugar_assay`` gives the drive of the benchmark's sugar trace
  (scripts/twin_audit.mjs:159-166);
- ``load_cord`` / ``cord_brain`` build the 21,843-cell ventral-cord brain of
  scripts/cord_kernel_ref.mjs:28-105 from cord_ir.bin, with its drive and its
  hemilineage ablations.

The data directory is the reference's ``public/data`` (and ``public/vision``
beside it for the eye map); nothing here is needed by the engine itself, which
takes the baked arrays.
"""
from __future__ import annotations

import json
import math
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
            "hl": at("hl"), "edgeClass": at("edgeClass"), "legOrder": list(d.get("legOrder", [])), "sets": sets}


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


# ------------------------------------------------------------------ the whole calibrated brain

def _type_names(conn: dict) -> list:
    """meta.types with JavaScript's reading of a missing name (``String(null)``)."""
    return ["null" if t is None else str(t) for t in conn["meta"]["types"]]


def by_type(conn: dict, *names) -> np.ndarray:
    """``D.byType(t)`` for each name in turn, concatenated (scripts/lib_node.mjs:12):
    the neurons of that exact type in ascending index order."""
    types = _type_names(conn)
    out = []
    for t in names:
        out.extend(i for i, x in enumerate(types) if x == t)
    return np.array(out, dtype=np.int64)


def _type_matcher(key: str):
    """A typeGain / classGain key: an exact name, or a regular expression source
    matched against the whole name when it holds a metacharacter
    (src/brainmodel.js:73-74)."""
    if re.search(r"[\\^$*+?()\[\]{}|]", key):
        rx = re.compile(f"^(?:{key})$")
        return lambda name: rx.match(name) is not None
    return lambda name: name == key


def _mulberry32(a: int):
    """JavaScript's mulberry32 on a 32-bit state, as src/brainmodel.js:85 writes it."""
    st = [a & 0xFFFFFFFF]

    def imul(x, y):
        return (x * y) & 0xFFFFFFFF

    def rnd():
        st[0] = (st[0] + 0x6D2B79F5) & 0xFFFFFFFF
        a0 = st[0]
        t = imul(a0 ^ (a0 >> 15), 1 | a0)
        t = ((t + imul(t ^ (t >> 7), 61 | t)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    return rnd


def type_gains(conn: dict, o: dict):
    """Per-neuron outgoing gain f32[N], or None when no gain is asked for
    (src/brainmodel.js:63-100).  Every multiply is stored to f32 as the
    Float32Array does.  ``typeGain`` multiplies the neurons of matching types,
    then ``neuronGain`` {sigma, seed} draws a lognormal factor per neuron
    (mulberry32, Box-Muller), then ``classGain`` multiplies by annotated
    class, then ``neuronGainTable`` multiplies by exp(logGain).  The
    transcendental functions are numpy's; JavaScript's may differ in the last
    bit of the binary64 value before the f32 rounding."""
    spec, cspec = o.get("typeGain"), o.get("classGain")
    keys = [k for k in spec if spec[k] != 1] if spec else []
    ckeys = [k for k in cspec if cspec[k] != 1] if cspec else []
    table = o.get("neuronGainTable")
    ng = o.get("neuronGain")
    if not keys and not ckeys and not (ng and ng.get("sigma")) and table is None:
        return None
    N = int(conn["N"])
    types = _type_names(conn)
    g = np.ones(N, dtype=np.float32)
    for k in keys:
        m = _type_matcher(k)
        f = float(spec[k])
        for i in range(N):
            if m(types[i]):
                g[i] = np.float32(float(g[i]) * f)
    if ng and ng.get("sigma"):
        sigma, seed = float(ng["sigma"]), int(ng.get("seed", 1))
        rnd = _mulberry32((seed * 2654435761) & 0xFFFFFFFF)
        for i in range(N):
            u = rnd()
            v = rnd()
            if u <= 0:
                u = 1e-9
            g[i] = np.float32(float(g[i]) * math.exp(sigma * math.sqrt(-2 * math.log(u)) * math.cos(2 * math.pi * v)))
    if ckeys:
        classes = conn["meta"]["classes"]
        cls = np.asarray(conn["cls"], dtype=np.int64)
        for i in range(N):
            v = cspec.get(classes[cls[i]])
            if v is not None:
                g[i] = np.float32(float(g[i]) * float(v))
    if table is not None:
        if isinstance(table, dict):
            items = ((int(k), float(v)) for k, v in table.items())
        else:
            items = ((i, float(v)) for i, v in enumerate(table) if v)
        for i, v in items:
            g[i] = np.float32(float(g[i]) * math.exp(v))
    return g


def class_physiology(conn: dict, o: dict):
    """src/brainmodel.js:22-30: Kenyon cells get threshold offset ``kcThreshold``
    and the lamina monopolar cells L1-L5 a tonic bias ``laminaBias``.  Returns
    (thr f32[N], bias f32[N]) starting from zeros, as a new brain has them."""
    N = int(conn["N"])
    classes = conn["meta"]["classes"]
    cls = np.asarray(conn["cls"], dtype=np.int64)
    kc = np.array([classes[c] == "Kenyon_Cell" for c in cls])
    types = _type_names(conn)
    lam = np.array([re.fullmatch(r"L[1-5]", t) is not None for t in types])
    thr = np.zeros(N, dtype=np.float32)
    thr[kc] = np.float32(o["kcThreshold"])
    bias = np.zeros(N, dtype=np.float32)
    bias[lam] = np.float32(o["laminaBias"])
    return thr, bias


# src/sim/neuromod.js:27-29: the constants the fed tone reads
NEUROMOD_FED = {"oaFed": 2.0, "ipcFed": 5.0, "sfa": 0.5, "targetShift": 2.0, "targetHalf": 400.0}


def neuromod_tone(conn: dict, thr, min_syn: float, calib: dict | None = None) -> np.ndarray:
    """Thresholds after ``new Neuromod(data, brain, {minSyn}).modulate()`` with
    the shipped operator ('thrField'), as scripts/calib_eval.mjs:50 runs it
    (src/sim/neuromod.js:60-160):
    - the modulatory cells (octopaminergic types ``OA-*``, then ``IPC``) get
      ``thr = base + sfa x rate`` with the fed rates (2 Hz OA, 5 Hz IPC), base
      being the calibration's resting threshold or the current one;
    - every other target of an OA neuron's synapses (count >= minSyn) gets
      ``thr0 - targetShift x s / (s + targetHalf)``, s = sum over OA sources of
      count x 2 Hz.
    Float32Array stores round every stored value to f32; the arithmetic is
    binary64 as in JavaScript.  Returns a new thr array."""
    P = NEUROMOD_FED
    thr = np.asarray(thr, dtype=np.float32).copy()
    types = _type_names(conn)
    N = int(conn["N"])
    oa = [i for i in range(N) if types[i].startswith("OA-")]
    ipc = [i for i in range(N) if types[i] == "IPC"]
    cells = oa + ipc
    cal = {}
    if calib:
        cal.update({int(k): v for k, v in calib.get("oaThr", {}).items()})
        cal.update({int(k): v for k, v in calib.get("ipcThr", {}).items()})
    base = np.array([cal.get(i, thr[i]) for i in cells], dtype=np.float32)
    rate = np.array([P["oaFed"]] * len(oa) + [P["ipcFed"]] * len(ipc), dtype=np.float32)
    ip = np.asarray(conn["indptr"], dtype=np.int64)
    ix = np.asarray(conn["indices"], dtype=np.int64)
    w = np.asarray(conn["weights"])
    # targets in first-seen order (their order only matters for the exact
    # integer sums below, so any order gives the same result)
    pos: dict = {}
    src_t, src_w = [], []
    for j in oa:
        a, b = ip[j], ip[j + 1]
        keep = w[a:b] >= min_syn
        for q, c in zip(ix[a:b][keep].tolist(), w[a:b][keep].tolist()):
            if q not in pos:
                pos[q] = len(pos)
            src_t.append(pos[q])
            src_w.append(c)
    targets = np.array(list(pos.keys()), dtype=np.int64)
    thr0 = thr[targets].copy()
    # setCellThr at construction: shift is still zero
    for n, i in enumerate(cells):
        thr[i] = np.float32(float(base[n]) + P["sfa"] * float(rate[n]))
    # modulate(): x[t] += count x c in f32; every term is an integer, so exact
    x = np.zeros(len(targets), dtype=np.float64)
    np.add.at(x, np.array(src_t, dtype=np.int64), np.array(src_w, dtype=np.float64) * float(P["oaFed"]))
    x = x.astype(np.float32).astype(np.float64)
    shift = (P["targetShift"] * x / (x + P["targetHalf"])).astype(np.float32)
    is_cell = np.isin(targets, np.array(cells, dtype=np.int64))
    keep = ~is_cell
    thr[targets[keep]] = (thr0[keep].astype(np.float64) - shift[keep].astype(np.float64)).astype(np.float32)
    return thr


def odor_receptors(conn_or_bodymap) -> np.ndarray:
    """``ORN_ALL``: every olfactory receptor neuron of the body map, sensor by
    sensor (scripts/calib_eval.mjs:55)."""
    bm = conn_or_bodymap.get("bodymap", conn_or_bodymap)
    return np.array([i for s in bm["sensors"] if s["kind"] == "odor" for i in s["idx"]], dtype=np.int64)


def flyvis_coupled(vision_dir: str) -> np.ndarray:
    """The connectome neurons coupled to the eye model, left eye then right
    (public/vision/flyvis_map.json); makeBrain masks their synaptic input as
    sensory (scripts/calib_eval.mjs:44)."""
    with open(os.path.join(vision_dir, "flyvis_map.json")) as f:
        fv = json.load(f)
    return np.array([p[0] for sd in ("L", "R") for p in fv["eyes"][sd]["pairs"]], dtype=np.int64)


def calibrated_brain(data_dir: str, vision_dir: str | None = None, conn: dict | None = None) -> dict:
    """The whole calibrated brain as scripts/calib_eval.mjs:43-51 makes it:
    calibrated parameters; input scale and sensory mask (widened with the
    eye-coupled neurons); octopaminergic fast synapses removed; weights baked
    by writeGraph with the type gains; then class physiology and the fed
    octopamine tone on thresholds and bias.  Returns the baked arrays and the
    start values of thr and bias for a new brain of any seed."""
    p = calibrated_params(data_dir)
    vision_dir = vision_dir or os.path.join(os.path.dirname(os.path.abspath(data_dir)), "vision")
    conn = conn or load_connectome(data_dir)
    with open(os.path.join(data_dir, "bodymap.json")) as f:
        conn["bodymap"] = json.load(f)
    in_scale, sensory, _ = brain_scales(conn, conn["size"], p)
    sensory = sensory.copy()
    sensory[flyvis_coupled(vision_dir)] = 1
    pre_sign = modulatory_sign(conn, conn["ntsign"], p)
    W, sign = write_graph(conn["N"], conn["indptr"], conn["indices"], conn["weights"], pre_sign, p, in_scale,
                          sensory, type_gains(conn, p))
    thr, bias = class_physiology(conn, p)
    if p.get("neuromod"):
        thr = neuromod_tone(conn, thr, p["minSyn"])
    return {"N": conn["N"], "indptr": conn["indptr"], "indices": conn["indices"], "W": W, "sign": sign,
            "thr": thr, "bias": bias, "params": p, "conn": conn}


def sugar_assay(conn: dict) -> dict:
    """The sugar trace of scripts/twin_audit.mjs:159-166: the olfactory receptors
    at 6 Hz throughout, the sugar gustatory neurons (types LB3b, LB3c) at 100 Hz
    for steps 0-799 and off from step 800, 1,400 steps of dt 0.5 ms."""
    return {"orn": odor_receptors(conn), "sugar": by_type(conn, "LB3b", "LB3c"), "orn_hz": 6.0,
            "sugar_hz": 100.0, "on_steps": 800, "steps": 1400}


def encode_targets(indptr, indices) -> np.ndarray:
    """Row targets (ascending within each row) as gaps in four byte planes, a
    form that compresses to about a third of the raw indices."""
    ip = np.asarray(indptr, dtype=np.int64)
    ix = np.asarray(indices, dtype=np.int64)
    first = np.zeros(ix.size, dtype=bool)
    first[ip[:-1][np.diff(ip) > 0]] = True
    prev = np.concatenate([[0], ix[:-1]])
    gaps = np.where(first, ix, ix - prev - 1)
    if (gaps < 0).any():
        raise ValueError("rows are not in ascending target order")
    g = gaps.astype(np.uint32)
    return np.stack([(g >> (8 * k)) & 255 for k in range(4)]).astype(np.uint8)


def decode_targets(indptr, planes) -> np.ndarray:
    """Inverse of the compact target coding: gaps in four byte planes -> row
    targets (the first entry of a row is absolute, later ones previous + gap + 1)."""
    ip = np.asarray(indptr, dtype=np.int64)
    g = np.zeros(planes.shape[1], dtype=np.int64)
    for k in range(4):
        g |= planes[k].astype(np.int64) << (8 * k)
    starts = ip[:-1][np.diff(ip) > 0]
    step = g + 1
    step[starts] = g[starts]
    cs = np.cumsum(step)
    # entry k of a row is the sum of the steps from the row's first entry to k
    before = cs[starts] - step[starts]
    off = np.zeros(g.size, dtype=np.int64)
    off[starts] = np.diff(np.concatenate([[0], before]))
    return (cs - np.cumsum(off)).astype(np.uint32)


def save_compact(path: str, data_dir: str, vision_dir: str | None = None, extra: dict | None = None):
    """Writes the calibrated brain to ``path`` in a compact form that needs no
    data directory to load: synapse counts, row targets as gap planes and the
    per-neuron inputs of writeGraph (sign, input scale, sensory mask), with the
    start thresholds and bias.  ``extra`` arrays are stored beside them."""
    p = calibrated_params(data_dir)
    vision_dir = vision_dir or os.path.join(os.path.dirname(os.path.abspath(data_dir)), "vision")
    conn = load_connectome(data_dir)
    b = calibrated_brain(data_dir, vision_dir, conn=conn)
    if type_gains(conn, p) is not None:
        raise ValueError("the compact form holds no type gains")
    in_scale, sensory, _ = brain_scales(conn, conn["size"], p)
    sensory = sensory.copy()
    sensory[flyvis_coupled(vision_dir)] = 1
    pre_sign = np.asarray(modulatory_sign(conn, conn["ntsign"], p), dtype=np.float32)
    params = {k: v for k, v in p.items() if isinstance(v, (int, float, bool))}
    np.savez_compressed(path, params=np.array(json.dumps(params)), indptr=np.asarray(conn["indptr"], dtype=np.uint32),
                        target_planes=encode_targets(conn["indptr"], conn["indices"]),
                        counts=np.asarray(conn["weights"], dtype=np.uint16), pre_sign=pre_sign,
                        in_scale=in_scale.astype(np.float32), sensory=sensory.astype(np.uint8), thr=b["thr"],
                        bias=b["bias"], **(extra or {}))
    return conn


def load_compact(path: str) -> dict:
    """Reads ``save_compact``'s file and bakes the weights again (writeGraph's
    products are deterministic, so W and sign come out bit-identical)."""
    z = np.load(path, allow_pickle=False)
    p = json.loads(str(z["params"]))
    indptr = z["indptr"]
    N = int(indptr.size - 1)
    indices = decode_targets(indptr, z["target_planes"])
    W, sign = write_graph(N, indptr, indices, z["counts"], z["pre_sign"], p, z["in_scale"], z["sensory"])
    out = {"N": N, "indptr": indptr, "indices": indices, "W": W, "sign": sign, "params": p}
    for k in z.files:
        if k not in ("params", "indptr", "target_planes", "counts", "pre_sign", "in_scale", "sensory"):
            out[k] = z[k]
    return out
