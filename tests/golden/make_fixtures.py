"""Copies a few of the reference's 3D scene fixtures and the OBJ meshes they
load into tests/golden/fixtures/ (scene files and meshes are data; the
scene files are rewritten with their mesh paths relative to this folder).

    python tests/golden/make_fixtures.py [/path/to/reference]
"""
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")
SCENES = [
    "fixtures/3D/unit-tests/5-cubes.json",
    "fixtures/3D/unit-tests/tet-corner.json",
    "fixtures/3D/unit-tests/edge-edge.json",
    "fixtures/3D/unit-tests/face-vertex.json",
    "fixtures/3D/friction/sliding.json",
    "fixtures/3D/examples/concave.json",
    "fixtures/3D/examples/kinematic.json",
    "fixtures/3D/examples/kinematically-scripted.json",
]

os.makedirs(os.path.join(OUT, "meshes"), exist_ok=True)
for rel in SCENES:
    with open(os.path.join(REF, rel)) as f:
        scene = json.load(f)
    for rb in scene.get("rigid_body_problem", {}).get("rigid_bodies", []):
        if rb.get("mesh"):
            src = os.path.join(REF, "meshes", rb["mesh"])
            dst = os.path.join(OUT, "meshes", os.path.basename(rb["mesh"]))
            shutil.copyfile(src, dst)
            rb["mesh"] = "meshes/" + os.path.basename(rb["mesh"])
    name = rel.split("/")[-2] + "-" + os.path.basename(rel)
    with open(os.path.join(OUT, name), "w") as f:
        json.dump(scene, f, indent=1)
    print(name)
