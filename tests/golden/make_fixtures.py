"""Writes tests/golden/cloth_armadillo_fixtures.npz from the reference's own
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
