"""Writes tests/golden/acquisition_3shots.json: the reference demo's acquisition
file (FWILab2d/FWILab2dCLangV1.2.6/demo/marm-ii/model_geometry/acquisition.json,
Marmousi-II, 30 shots x 601 receivers) cut to its first three shots, keeping
the "model" and "acquisition" blocks.  Data only; run once from the repository
root where the reference is present:

    python tests/golden/make_acquisition_fixture.py
"""
import json
import os

SRC = ("/root/reference/FWILab2d/FWILab2dCLangV1.2.6/demo/marm-ii/model_geometry/"
       "acquisition.json")
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "acquisition_3shots.json")

with open(SRC) as f:
    d = json.load(f)
d["record"] = d["record"][:3]
with open(DST, "w") as f:
    json.dump(d, f, indent=1)
print(f"wrote {DST}: {len(d['record'])} shots")
