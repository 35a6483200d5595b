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
            "fvGet", "fvBackward",
            "diffCreate", "diffDestroy", "diffSetParams", "diffSetInput", "diffSetEdgeGain", "diffSetMasks",
            "diffSetDriven", "diffSetLoss", "diffSetReadout", "diffSetTapeCapacity", "diffSetDriveSchedule",
            "diffSetLossPerEpoch", "diffDriveGrad", "diffSetSpikeTrain", "diffForward", "diffBackward", "diffTypeGrads", "diffGetState",
            "diffSpikes", "diffRng", "diffStat", "diffSize", "diffParamsLength",
            "diffBatchCreate", "diffBatchDestroy", "diffBatchForward", "diffBatchBackward", "diffBatchStat",
            "batchCreate", "batchDestroy", "batchSetDrive", "batchSetMember", "batchSetSteps", "batchRun",
            "batchGetArray", "batchSetLanes", "batchSetRng", "batchSetPoolFeedback", "batchPoolCounts", "batchRng",
            "batchCounters", "batchSetEngine", "batchStat", "batchSize"}
    assert want <= set(res["exports"])
    assert res["paramsLength"] == 19


@pytest.mark.parametrize("shim", ["lifhip.js", "flyviship.js", "lifdiffhip.js", "lifbatchhip.js"])
def test_shim_is_node12_syntax(shim):
    """The reference's Node is v12: no optional chaining or nullish coalescing."""
    src = open(os.path.join(ROOT, "fly-brain_amd", "js", shim)).read()
    assert "?." not in src and "??" not in src
    subprocess.run(["node", "--check", os.path.join(ROOT, "fly-brain_amd", "js", shim)], check=True, timeout=60)


def test_diff_params_vector_matches_the_header():
    """lifdiffhip.js packs fb_diff_params as 23 doubles then soft, checkpoint
    and exact32; the header's struct has those fields in that order."""
    import re
    txt = open(os.path.join(ROOT, "include", "flybrain.h")).read()
    end = txt.index("} fb_diff_params;")
    body = txt[txt.rindex("typedef struct {", 0, end):end]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    doubles = re.findall(r"double ([^;]+);", body)
    names = [n.strip() for d in doubles for n in d.split(",")]
    ints = [n.strip() for d in re.findall(r"int32_t ([^;]+);", body) for n in d.split(",")]
    js = open(os.path.join(ROOT, "fly-brain_amd", "js", "lifdiffhip.js")).read()
    fields = re.search(r"const DIFF_FIELDS = \[([^\]]+)\]", js).group(1)
    assert [f.strip().strip("'") for f in fields.replace("\n", " ").split(",")] == names
    assert ints == ["soft", "checkpoint", "exact32"]
