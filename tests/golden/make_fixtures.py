"""Writes tests/golden/cloth_armadillo_fixtures.npz (and ..._demo_rest.npz) from the reference's own
data files (example/ClothArmadillo/cloth{1,2}): the cloth meshes as the
reference's OBJ loader reads them (float32 vertices, 0-based triangles) and the
cached boundary.dat / patchInfo.dat that the reference computed from them
(formats: PdIpcModel.h:219-251 and :291-318).  Run in the build container only;
the tests read the .npz."""

import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example/ClothArmadillo"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cloth_armadillo_fixtures.npz")


def load_obj(path):
    V, F = [], []
    for line in open(path):
        if line.startswith("v "):
            V.append([np.float32(t) for t in line.split()[1:4]])
        elif line.startswith("f "):
            F.append([int(t.split("/")[0]) - 1 for t in line.split()[1:4]])
    return np.array(V, np.float32), np.array(F, np.int32)


arrays = {}
for name in ("cloth1", "cloth2"):
    d = os.path.join(REF, name)
    v, f = load_obj(os.path.join(d, name + ".obj"))
    arrays[name + "_v"] = v
    arrays[name + "_f"] = f
    arrays[name + "_boundary"] = np.fromfile(os.path.join(d, "boundary.dat"), np.int32)
    arrays[name + "_patch"] = np.fromfile(os.path.join(d, "patchInfo.dat"), np.int32)
    fx = open(os.path.join(d, "fixed_points.txt")).read().split()
    arrays[name + "_fixed"] = np.array([int(t) for t in fx[2:]], np.int32)
np.savez_compressed(OUT, **arrays)
print(OUT, {k: v.shape for k, v in arrays.items()})

# The rest of the shipped demo as data: config.xml (bytes) and the armadillo's
# surface OBJ, tet nodes / elements (0-based), fixed ids and surface embedding,
# so that tests can rebuild the demo's scene directory without the reference.
demo = {"config_xml": np.frombuffer(open(os.path.join(REF, "config.xml"), "rb").read(), np.uint8)}
d = os.path.join(REF, "armadillo")
v, f = load_obj(os.path.join(d, "armadillo.obj"))
demo["armadillo_v"], demo["armadillo_f"] = v, f
def records(path):  # the header's count of records (a trailing "# ..." line is not one)
    lines = [ln.split() for ln in open(path)]
    return lines[1:1 + int(lines[0][0])]


rows = records(os.path.join(d, "armadillo.node"))
assert int(rows[0][0]) == 1  # 1-based
demo["armadillo_nodes"] = np.array([[np.float32(t) for t in r[1:4]] for r in rows], np.float32)
rows = records(os.path.join(d, "armadillo.ele"))
demo["armadillo_tets"] = np.array([[int(t) - 1 for t in r[1:5]] for r in rows], np.int32)
fx = open(os.path.join(d, "fixed_points.txt")).read().split()
demo["armadillo_fixed"] = np.array([int(t) for t in fx[2:]], np.int32)
rows = [ln.split() for ln in open(os.path.join(d, "tetMeshInterpolation.txt"))][1:]
demo["armadillo_interp_eid"] = np.array([int(r[1]) for r in rows if r], np.int32)
demo["armadillo_interp_w"] = np.array([[np.float32(t) for t in r[2:6]] for r in rows if r], np.float32)
OUT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cloth_armadillo_demo_rest.npz")
np.savez_compressed(OUT2, **demo)
print(OUT2, {k: v.shape for k, v in demo.items()})
