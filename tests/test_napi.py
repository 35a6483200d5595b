"""The Node binding (N-API addon + LIFHip class): the addon loads and exports
the brain API the JS shim needs.  No device calls; the GPU parity of the same
path is in test_gpu_napi.py."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ADDON = os.path.join(ROOT, "fly-brain_amd", "lib", "flybrain_napi.node")

pytestmark = pytest.mark.skipif(not shutil.which("node") or not os.path.exists(ADDON),
                                reason="node or the built addon is absent")


def test_addon_exports_the_brain_api(tmp_path):
    spec = tmp_path / "spec.json"
    spec.write_text(json.dumps({"exportsOnly": True}))
    out = subprocess.run(["node", os.path.join(HERE, "js", "napi_run.js"), str(spec)], capture_output=True,
                         text=True, timeout=60, check=True)
    res = json.loads(out.stdout)
    want = {"graphCreate", "graphRelease", "brainCreate", "brainDestroy", "reset", "setParams", "setDrive",
            "setDriveOne", "setBias", "setThr", "addG", "pulse", "setBackground", "step", "fired", "getArray",
            "setArray", "scalars", "counters", "runAsync", "deviceInfo", "paramsLength", "applyDeltas",
            "fvModelCreate", "fvModelSetKdt", "fvModelSetInputMap", "fvModelRelease", "fvCreate", "fvDestroy",
            "fvReset", "fvSetInput", "fvSetLum", "fvSetV", "fvStep", "fvTapeStart", "fvStepTaped", "fvTapeLen",
            "fvGet", "fvBackward"}
    assert want <= set(res["exports"])
    assert res["paramsLength"] == 19


def test_shim_is_node12_syntax():
    """The reference's Node is v12: no optional chaining or nullish coalescing."""
    src = open(os.path.join(ROOT, "fly-brain_amd", "js", "lifhip.js")).read()
    assert "?." not in src and "??" not in src
    subprocess.run(["node", "--check", os.path.join(ROOT, "fly-brain_amd", "js", "lifhip.js")], check=True,
                   timeout=60)
