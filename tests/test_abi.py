"""The C-ABI library loads and exports every symbol include/flybrain.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import importlib
import os

import pytest

L = importlib.import_module("fly-brain_amd.lib")


def test_library_is_built():
    assert os.path.exists(L.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_header_symbol():
    lib = L.load()
    syms = L.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(L.header_symbols()) <= set(L._SIGS), set(L.header_symbols()) - set(L._SIGS)


def test_last_error_is_empty_string_initially():
    lib = L.load()
    assert isinstance(lib.fb_last_error(), bytes)


def test_argument_errors_do_not_touch_the_device():
    lib = L.load()
    out = ctypes.c_void_p()
    # N <= 0 is rejected before any HIP call
    rc = lib.fb_graph_create(0, 0, None, None, None, None, None, ctypes.byref(out))
    assert rc == 1
    assert b"bad arguments" in lib.fb_last_error()


def test_params_struct_matches_header():
    P = importlib.import_module("fly-brain_amd.params")
    txt = open(L.HEADER).read()
    end = txt.index("} fb_params;")
    body = txt[txt.rindex("typedef struct {", 0, end):end]
    import re
    fields = re.findall(r"\b([A-Za-z]+)\s*[,;]", body.replace("double", ""))
    assert tuple(fields) == P.PARAM_FIELDS
    assert ctypes.sizeof(P.FbParams) == 8 * 19
