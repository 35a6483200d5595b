"""ctypes binding of libflybrain.so (the C ABI in include/flybrain.h).

The library is loaded from fly-brain_amd/lib/ only.  There is no fallback:
if it is missing, loading raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libflybrain.so")
HEADER = os.path.join(ROOT, "include", "flybrain.h")

FB_OK = 0
ENGINES = {"auto": 0, "launch": 1, "graph": 2, "persistent": 3}
ARRAYS = ("v", "gE", "gI", "refr", "trace", "adapt", "res", "bias", "thr", "drive", "spikeCount")

_lib = None


class FlyBrainError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"flybrain error {code}: {msg}")
        self.code = code


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
D = ctypes.c_double
F = ctypes.c_float
PP = ctypes.POINTER(ctypes.c_void_p)
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)

_SIGS = {
    "fb_last_error": (ctypes.c_char_p, []),
    "fb_device_info": (I32, [ctypes.c_char_p, I32, PI32]),
    "fb_graph_create": (I32, [I32, I64, P, P, P, P, P, PP]),
    "fb_graph_retain": (I32, [P]),
    "fb_graph_release": (I32, [P]),
    "fb_brain_create": (I32, [P, P, D, PP]),
    "fb_brain_destroy": (I32, [P]),
    "fb_reset": (I32, [P]),
    "fb_set_params": (I32, [P, P]),
    "fb_get_params": (I32, [P, P]),
    "fb_set_drive": (I32, [P, P, I32, D]),
    "fb_set_drive_one": (I32, [P, I32, D]),
    "fb_set_bias": (I32, [P, P, I32, D]),
    "fb_set_thr": (I32, [P, I32, D]),
    "fb_add_g": (I32, [P, I32, D, D]),
    "fb_apply_deltas": (I32, [P, I32, P, P, P]),
    "fb_apply_deltas_noalloc": (I32, [P, I32, P, P, P]),
    "fb_pulse": (I32, [P, P, I32, D]),
    "fb_set_background": (I32, [P, D, D]),
    "fb_driven": (I32, [P, P, I32, PI32]),
    "fb_step": (I32, [P, I32, I32, PI32]),
    "fb_fired": (I32, [P, P, I32, PI32]),
    "fb_run_record": (I32, [P, I32, I32, P, P, I64, PI64]),
    "fb_get_array": (I32, [P, I32, P]),
    "fb_set_array": (I32, [P, I32, P]),
    "fb_brain_scalars": (I32, [P, ctypes.POINTER(U32), PI32, ctypes.POINTER(F), ctypes.POINTER(D)]),
    "fb_brain_size": (I32, [P, PI32]),
    "fb_brain_counters": (I32, [P, PI64, PI64, PI64, PI64]),
    "fb_brain_set_stream": (I32, [P, P]),
    "fb_brain_stat": (I32, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
    "fb_batch_create": (I32, [P, P, I32, P, P, I32, P, PP]),
    "fb_batch_destroy": (I32, [P]),
    "fb_batch_set_drive": (I32, [P, P, I32, D]),
    "fb_batch_set_member": (I32, [P, I32, I32, P, I32, D]),
    "fb_batch_set_steps": (I32, [P, P]),
    "fb_batch_run": (I32, [P, I32, P, P, P, P, I64, PI64]),
    "fb_batch_get_array": (I32, [P, I32, I32, P]),
    "fb_batch_rng": (I32, [P, P]),
    "fb_batch_size": (I32, [P, PI32, PI32, PI32]),
    "fb_batch_counters": (I32, [P, PI64, PI64, PI64, PI64]),
    "fb_batch_set_stream": (I32, [P, P]),
    "fb_batch_set_engine": (I32, [P, I32]),
    "fb_batch_set_shape": (I32, [P, I32]),
    "fb_batch_stat": (I32, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
    "fb_diff_create": (I32, [I32, I64, P, P, P, P, P, PP]),
    "fb_diff_destroy": (I32, [P]),
    "fb_diff_set_params": (I32, [P, P]),
    "fb_diff_get_params": (I32, [P, P]),
    "fb_diff_set_input": (I32, [P, I32, P]),
    "fb_diff_set_edge_gain": (I32, [P, P]),
    "fb_diff_set_masks": (I32, [P, P, P]),
    "fb_diff_set_driven": (I32, [P, P, I32]),
    "fb_diff_set_loss": (I32, [P, P, I32]),
    "fb_diff_set_readout": (I32, [P, P, I32, P, D]),
    "fb_diff_set_tape_capacity": (I32, [P, I64]),
    "fb_diff_set_drive_schedule": (I32, [P, P, I32, I32, P, I32, I32]),
    "fb_diff_set_loss_per_epoch": (I32, [P, P, I32, I32]),
    "fb_diff_drive_grad": (I32, [P, P]),
    "fb_diff_set_spike_train": (I32, [P, I32, P, P]),
    "fb_set_rng": (I32, [P, ctypes.c_uint32]),
    "fb_batch_set_lanes": (I32, [P, I32, P]),
    "fb_batch_set_rng": (I32, [P, P]),
    "fb_batch_set_pool_feedback": (I32, [P, P, I32, P, I32, ctypes.c_float, ctypes.c_float, P, I32]),
    "fb_batch_pool_counts": (I32, [P, P, I32, P]),
    "fb_set_pool_feedback": (I32, [P, P, I32, P, I32, ctypes.c_float, ctypes.c_float, I32, I32]),
    "fb_pool_counts": (I32, [P, P, I32, P]),
    "fb_fv_model_create": (I32, [I32, I64, P, P, P, P, P, D, P]),
    "fb_fv_model_release": (I32, [P]),
    "fb_fv_create": (I32, [P, P]),
    "fb_fv_destroy": (I32, [P]),
    "fb_fv_set_input": (I32, [P, P]),
    "fb_fv_step": (I32, [P, I32]),
    "fb_fv_get": (I32, [P, I32, P]),
    "fb_fv_set_v": (I32, [P, P]),
    "fb_fv_reset": (I32, [P]),
    "fb_fv_model_set_kdt": (I32, [P, P]),
    "fb_fv_model_set_input_map": (I32, [P, I32, P, P]),
    "fb_fv_set_lum": (I32, [P, P]),
    "fb_fv_tape_start": (I32, [P, I32]),
    "fb_fv_step_taped": (I32, [P, I32]),
    "fb_fv_tape_len": (I32, [P, P]),
    "fb_fv_size": (I32, [P, PI32, P, PI32]),
    "fb_fv_model_size": (I32, [P, PI32, P]),
    "fb_fv_backward": (I32, [P, P, P, P, P, P]),
    "fb_diff_forward": (I32, [P, I32, D, I32, ctypes.POINTER(D)]),
    "fb_diff_backward": (I32, [P, I32, P, P, P, P, P]),
    "fb_diff_type_grads": (I32, [P, P, I32, P]),
    "fb_diff_get_state": (I32, [P, I32, P]),
    "fb_diff_spikes": (I32, [P, P, P, P, I64, PI64]),
    "fb_diff_readout": (I32, [P, P, P]),
    "fb_diff_rng": (I32, [P, P]),
    "fb_diff_size": (I32, [P, PI32, P, PI32, PI32]),
    "fb_diff_set_stream": (I32, [P, P]),
    "fb_diff_stat": (I32, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
    "fb_diff_grads": (I32, [P, P, P, P, P, P]),
    "fb_diff_batch_create": (I32, [P, I32, PP]),
    "fb_diff_batch_destroy": (I32, [P]),
    "fb_diff_batch_forward": (I32, [P, P, P, I32, P]),
    "fb_diff_batch_backward": (I32, [P, I32]),
    "fb_diff_batch_stat": (I32, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
    "fb_synchronize": (I32, [P]),
}


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    txt = open(HEADER).read()
    pat = r"^\s*(?:const\s+)?[A-Za-z_0-9]+\s*\*?\s*(fb_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(pat, txt, flags=re.MULTILINE)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlyBrainError(-1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int):
    if code != FB_OK:
        msg = load().fb_last_error().decode(errors="replace")
        raise FlyBrainError(code, msg)


def ptr(a) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)
