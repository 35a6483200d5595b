"""Copies the reference's flyvis model files (public/vision/flyvis.bin,
flyvis.json, flyvis_inputs.json) and the eye-to-CNS pair map
(flyvis_map.json) into tests/golden/_vision/ for the GPU tests and the bench
of the optic lobe and the visual chain, which cannot read the reference at
run time.

Run where the reference's data exist (`__graft_entry__.build()` does it):
    python tests/golden/make_vision_fixture.py [/root/reference/public/vision]
The copies are data, stay out of git (tests/golden/_vision/ is ignored) and
travel with the working tree to the GPU machine.
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ("flyvis.bin", "flyvis.json", "flyvis_inputs.json", "flyvis_map.json")


def main(src="/root/reference/public/vision"):
    if not all(os.path.exists(os.path.join(src, f)) for f in FILES):
        sys.exit(f"no flyvis model under {src}")
    out = os.path.join(HERE, "_vision")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(out, f))
    print("wrote", out)


if __name__ == "__main__":
    main(*sys.argv[1:])
