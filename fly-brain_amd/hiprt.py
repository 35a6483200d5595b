"""Minimal ctypes access to the HIP runtime: streams and events for timing.

Plumbing only (no compute): the benchmark times the engine's kernels with HIP
events recorded on the same stream the kernels are queued on.
"""
from __future__ import annotations

import ctypes
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = "libamdhip64.so"
        for cand in ("/opt/rocm/lib/libamdhip64.so", "/opt/rocm/lib/libamdhip64.so.7"):
            if os.path.exists(cand):
                path = cand
                break
        _hip = ctypes.CDLL(path)
        P = ctypes.c_void_p
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(P)]
        _hip.hipStreamDestroy.argtypes = [P]
        _hip.hipStreamSynchronize.argtypes = [P]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        _hip.hipEventDestroy.argtypes = [P]
        _hip.hipEventRecord.argtypes = [P, P]
        _hip.hipEventSynchronize.argtypes = [P]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class Stream:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.h)), "hipStreamCreate")

    @property
    def handle(self) -> int:
        return self.h.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        """Releases the stream; every handle queued on it must be gone first."""
        if self.h.value:
            _ok(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")
            _ok(hip().hipStreamDestroy(self.h), "hipStreamDestroy")
            self.h = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream: Stream):
        _ok(hip().hipEventRecord(self.h, stream.h), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def destroy(self):
        if self.h.value:
            _ok(hip().hipEventDestroy(self.h), "hipEventDestroy")
            self.h = ctypes.c_void_p()

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.h, later.h), "hipEventElapsedTime")
        return float(ms.value)


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
