"""Model parameters in the reference's vocabulary.

``DEFAULTS`` restates ``DEFAULTS`` of src/lif.js:6-15 (Lulzx/fly-brain).  The
engine takes the subset the kernel reads, as f64 values, in the field order of
``fb_params`` (include/flybrain.h); every derived constant is computed in f64
on the host and rounded once to f32, as src/lifwasm.js:58-67 does.
"""
from __future__ import annotations

import ctypes

DEFAULTS = {
    "dt": 0.5, "vRest": -52.0, "vThresh": -45.0, "vReset": -52.0, "tauM": 20.0,
    "tauSyn": 5.0, "tRef": 2.2, "delay": 1.8, "wSyn": 0.275, "noise": 0.0,
    "traceTau": 30.0, "adaptInc": 2.0, "adaptTau": 100.0, "depU": 0.2, "depTau": 200.0,
    "minSyn": 1, "preInh": 0.0, "coba": False, "eExc": 0.0, "eInh": -70.0, "inhGain": 1.0,
}

# fb_params field order (include/flybrain.h)
PARAM_FIELDS = ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay",
                "traceTau", "adaptInc", "adaptTau", "depU", "depTau", "coba", "eExc",
                "eInh", "bgRate", "bgAmp", "preInh")


class FbParams(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in PARAM_FIELDS]


def merged(params: dict | None = None) -> dict:
    """``{...DEFAULTS, ...params}`` as the JS constructors do."""
    p = dict(DEFAULTS)
    if params:
        p.update(params)
    return p


def to_struct(p: dict) -> FbParams:
    s = FbParams()
    for name in PARAM_FIELDS:
        val = p.get(name, 0.0)
        if val is None:
            val = 0.0
        setattr(s, name, float(bool(val)) if name == "coba" else float(val))
    return s


def config_params(coba: bool = False, **over) -> dict:
    """Constants of the synthetic workloads: dt 0.1 ms, tau_m 20 ms, tau_syn 5 ms,
    v_rest -52, v_th -45, v_reset -52 mV, refractory 2.2 ms, delay 1.8 ms, and
    adaptation / depression off (the calibrated model's values).  The COBA
    variant uses E_exc 0 mV and E_inh -80 mV."""
    p = merged({"dt": 0.1, "adaptInc": 0.0, "depU": 0.0, "coba": coba,
                "eExc": 0.0, "eInh": -80.0 if coba else -70.0})
    p.update(over)
    return p


def nslots(p: dict) -> int:
    """``max(1, Math.round(delay / dt)) + 1`` (src/lifwasm.js:39)."""
    import math
    x = p["delay"] / p["dt"]
    f = math.floor(x)
    r = f + 1 if x - f >= 0.5 else f
    return max(1, int(r)) + 1


# ---------------------------------------------------------------- differentiable model
#
# Options of LIFDiff (src/lifdiff.js:88-112), the model the adjoint runs.  Its
# defaults differ from lif.js's (wSyn 0.3, adaptInc 0, depU 0, COBA always,
# eInh -70, surrogateBeta 2).  `checkpoint` is the segment length K of the
# replay (the reference fixes it at 16, src/lifdiff.js:69).
DIFF_DEFAULTS = {
    "dt": 0.5, "vRest": -52.0, "vThresh": -45.0, "vReset": -52.0, "tauM": 20.0, "tauSyn": 5.0,
    "tRef": 2.2, "delay": 1.8, "wSyn": 0.3, "adaptInc": 0.0, "adaptTau": 100.0, "depU": 0.0,
    "depTau": 200.0, "eExc": 0.0, "eInh": -70.0, "inhGain": 1.0, "kcThreshold": 0.0,
    "laminaBias": 0.0, "surrogateBeta": 2.0, "adjClip": 0.0, "sizeAlpha": 0.0,
    "maxSizeScale": 20.0, "boostCap": 1.0, "soft": False, "checkpoint": 16, "minSyn": 1, "exact32": False,
}

# fb_diff_params field order (include/flybrain.h): doubles, then two int32
DIFF_FIELDS = ("dt", "vRest", "vThresh", "vReset", "tauM", "tauSyn", "tRef", "delay",
               "wSyn", "adaptInc", "adaptTau", "depU", "depTau",
               "eExc", "eInh", "inhGain", "kcThreshold", "laminaBias",
               "surrogateBeta", "adjClip", "sizeAlpha", "maxSizeScale", "boostCap")
DIFF_INT_FIELDS = ("soft", "checkpoint", "exact32")

# gradient order (src/lifdiff.js:75)
DIFF_PARAMS = ("wSyn", "sizeAlpha", "inhGain", "eInh", "vThresh", "kcThreshold", "laminaBias",
               "adaptInc", "depU")


class FbDiffParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in DIFF_FIELDS] + [(n, ctypes.c_int32) for n in DIFF_INT_FIELDS]


def diff_merged(params: dict | None = None) -> dict:
    p = dict(DIFF_DEFAULTS)
    if params:
        p.update(params)
    return p


def to_diff_struct(p: dict) -> FbDiffParams:
    s = FbDiffParams()
    for n in DIFF_FIELDS:
        setattr(s, n, float(p[n]))
    s.soft = 1 if p["soft"] else 0
    s.checkpoint = int(p["checkpoint"])
    s.exact32 = 1 if p.get("exact32") else 0
    return s


def gated_counts(counts, indices, min_syn=1, sensory=None):
    """The delivered synapse count per edge: zero where the count is under
    minSyn or the target is a sensory (externally driven) neuron, as
    LIFDiff's gate (src/lifdiff.js:167-172)."""
    import numpy as np
    c = np.asarray(counts, dtype=np.float32).copy()
    keep = c >= min_syn
    if sensory is not None:
        keep &= ~np.asarray(sensory, dtype=bool)[np.asarray(indices)]
    c[~keep] = 0.0
    return c
