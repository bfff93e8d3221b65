"""The reference simulator's scene files, read into a scene of this package (the
dict `pdipc.Simulator.from_scene` and the oracle take).

A scene directory holds `config.xml` and one folder per object
(`<name>/<name>.obj`, for volumetric objects also `<name>.node` / `<name>.ele`,
optionally `fixed_points.txt`, `tetMeshInterpolation.txt` and the cached
`boundary.dat` / `patchInfo.dat`).  Reading follows the reference:

* `config.xml`: the `<simulator>` element of `<SimFramework>`
  (BaseSimulator::readSimulatorFromConfigFile, Simulator/BaseSimulator.cpp:4-65):
  `timeStep`, `gravity`, and the PD-IPC attributes `dhat`, `kappa` (`tol`,
  `damping`, `substep`, `penalty` are read by the reference but not used by
  its step; they are kept in `scene["extra"]`); `<object>` and `<static_object>`
  children in file order.  Elements after `</SimFramework>` are not part of
  the simulator (TinyXML keeps them as further document children, the reader
  never visits them).
* an object (PdIpcSimulator::addModelFromConfigFile, PdIpcSimulator.h:462-526;
  BaseSimulator::getModelFilename / addModelFromFile,
  Simulator/BaseSimulator.cpp:72-200): `type` = model type (1 volumetric,
  otherwise cloth) and the stiffness triple; `patchDim`, `scale`,
  `translation`, `rotation` (degrees about x, y, z; the fourth value is read
  and unused), `density`; attributes `uniform` / `origin` (default on; "off",
  "false" or "0" turn them off) and `volumetric_mesh="tet"`.
* geometry in float32, the reference's `qeal` (DataCore.h:8-25), with its
  order of operations (Model/BaseModel.cpp:21-66, 405-475;
  Model/TriangleMesh/BaseSurfaceMesh.cpp:384-480; Model/VolumetricMesh/
  BaseTetMesh.cpp:368-410): on loading, `placeOrigin` (subtract the bounding-box
  centre) if `origin`, then `uniform` (divide by the largest extent) if
  `uniform`; then scale about the centre, translate, rotate about the centre.
  Box and centre come from the surface mesh (also for volumetric objects, whose
  tet nodes get the same transform), computed as `computeBBox` does: a point
  that lowers a minimum is not compared with the maximum.
* meshes: OBJ as tinyobjloader returns it with triangulation on (vertices in
  file order, faces of every group in order, polygons as fans, 1-based or
  negative indices); `.node` / `.ele` as BaseTetMesh::readNode/ElementFromTetFormat
  (Model/VolumetricMesh/BaseTetMesh.cpp:26-140: index base taken from the first
  record); `fixed_points.txt` as PdIpcModel::readFixedInfo
  (Simulator/PD-IPC/PdIpcModel.h:56-77: "c N" then N ids; the reference's
  `std::unique` result is discarded, so repeated ids stay).
* the surface of a volumetric object follows its tet mesh through
  `tetMeshInterpolation.txt` (BaseModel::computeTetMeshInterpolation and
  updateSurfaceFromTetMesh, Model/BaseModel.cpp:249-300, 361-397):
  `surface_positions` applies it.

Decimal strings become float32 through float64 (numpy), which equals the
reference's direct conversion except for strings within one float64 rounding
of a float32 rounding boundary.
"""

import os
import re
import xml.etree.ElementTree as ET

import numpy as np

import scenes

F32 = np.float32


# ---------------------------------------------------------------- file readers
def read_obj(path):
    """(vertices float32 (n, 3), triangles int32 (m, 3), 0-based)."""
    verts, faces = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                verts.append([float(t) for t in tok[1:4]])
            elif tok[0] == "f":
                n = len(verts)
                idx = []
                for t in tok[1:]:
                    i = int(t.split("/")[0])
                    idx.append(i - 1 if i > 0 else n + i)
                for k in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[k], idx[k + 1]])
    v = np.array(verts, dtype=np.float64).astype(F32).reshape(-1, 3)
    return v, np.array(faces, dtype=np.int32).reshape(-1, 3)


def _records(path):
    with open(path) as fh:
        head = fh.readline().split()
        rows = [ln.split() for ln in fh if ln.strip() and not ln.lstrip().startswith("#")]
    return head, rows


def read_tet(node_path, ele_path):
    """(nodes float32 (n, 3), tets int32 (t, 4), 0-based)."""
    head, rows = _records(node_path)
    n = int(head[0])
    base = 1 if int(rows[0][0]) == 1 else 0
    pts = np.zeros((n, 3), np.float64)
    for r in rows[:n]:
        pts[int(r[0]) - base] = [float(t) for t in r[1:4]]
    head, rows = _records(ele_path)
    t = int(head[0])
    base = 1 if int(rows[0][0]) == 1 else 0
    tets = np.zeros((t, 4), np.int32)
    for r in rows[:t]:
        tets[int(r[0]) - base] = [int(x) - base for x in r[1:5]]
    return pts.astype(F32), tets


def read_fixed(path):
    """The ids of fixed_points.txt ("c N" then N ids), as the reference keeps
    them: its std::unique result is discarded, so the vector is the unique
    prefix followed by the remaining original entries (its length unchanged)."""
    with open(path) as fh:
        tok = fh.read().split()
    if len(tok) < 2 or tok[0] != "c":
        return np.zeros(0, np.int32)
    ids = [int(t) for t in tok[2:2 + int(tok[1])]]
    out = []
    for i in ids:
        if not out or out[-1] != i:
            out.append(i)
    out += ids[len(out):]
    return np.array(out, np.int32)


def read_interpolation(path):
    """tetMeshInterpolation.txt: (tet id per surface vertex, 4 weights float32)."""
    with open(path) as fh:
        n = int(fh.readline().split()[0])
        eid = np.zeros(n, np.int32)
        w = np.zeros((n, 4), np.float64)
        for i in range(n):
            r = fh.readline().split()
            eid[i] = int(r[1])
            w[i] = [float(t) for t in r[2:6]]
    return eid, w.astype(F32)


# ------------------------------------------------------------ transforms (float32)
def bbox(points):
    """BaseSurfaceMesh::computeBBox: per axis, a point below the running minimum
    updates it and is not compared with the maximum (else-if)."""
    lo = np.full(3, np.finfo(F32).max, F32)
    hi = np.full(3, -np.finfo(F32).max, F32)
    for p in points:
        for k in range(3):
            if p[k] < lo[k]:
                lo[k] = p[k]
            elif p[k] > hi[k]:
                hi[k] = p[k]
    return lo, hi


def center(points):
    lo, hi = bbox(points)
    return ((lo + hi) / F32(2.0)).astype(F32)


def _rotation(rx, ry, rz):
    a = np.radians([rx, ry, rz])
    cx, sx, cy, sy, cz, sz = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1]), np.cos(a[2]), np.sin(a[2])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def transform(surface, nodes, uniform, origin, scale, translation, rotation):
    """The reference's load-time and configured transforms of one object, in
    float32; `nodes` (tet nodes or None) move with the surface."""
    s = surface.astype(F32).copy()
    t = None if nodes is None else nodes.astype(F32).copy()
    if origin:  # BaseModel::placeOrigin
        c = center(s)
        s -= c
        if t is not None:
            t -= c
    if uniform:  # BaseSurfaceMesh::uniform, BaseTetMesh::uniform(div); the box is
        # the current one (with `origin` off the reference reads a box member that
        # nothing has computed yet: indeterminate values)
        lo, hi = bbox(s)
        ext = hi - lo
        div = ext[0]
        if div < ext[1]:
            div = ext[1]
        if div < ext[2]:
            div = ext[2]
        if abs(div) < 1e-6:
            div = F32(1.0)
        s /= div
        if t is not None:
            t /= div
    sc = np.asarray(scale, F32)
    c = center(s)  # BaseModel::scaleModel: about the centre, by s / 1
    s = (s - c) * sc + c
    if t is not None:
        t = (t - c) * sc + c
    tr = np.asarray(translation, F32)  # BaseModel::translateModel: by t - 0
    s = s + tr
    if t is not None:
        t = t + tr
    rx, ry, rz = (float(r) if abs(float(r)) > 1e-6 else 0.0 for r in rotation[:3])
    if rx or ry or rz:  # BaseModel::rotateModel (non-zero rotations: parity unpinned)
        R = _rotation(rx, ry, rz)
        c = center(s).astype(np.float64)
        s = ((s.astype(np.float64) - c) @ R.T + c).astype(F32)
        if t is not None:
            t = ((t.astype(np.float64) - c) @ R.T + c).astype(F32)
    return s, t


# ----------------------------------------------------------------- config.xml
_PARAM_KEYS = {"timeStep": "dt", "gravity": "gravity", "dhat": "dhat", "kappa": "kappa"}
_UNUSED_KEYS = ("tol", "damping", "substep", "penalty")


def _floats(el):
    return [float(t) for t in (el.text or "").split()]


def _flag(value):
    return not (value in ("off", "false", "0"))


def load_config(path, **params):
    """The scene of a reference `config.xml` (path to the file or to its
    directory).  `params` override step parameters (e.g. lg_iters).  Volumetric
    objects keep their surface mesh and interpolation under `scene["surfaces"]`
    (see `surface_positions`)."""
    if os.path.isdir(path):
        path = os.path.join(path, "config.xml")
    root_dir = os.path.dirname(os.path.abspath(path))
    text = open(path).read()
    m = re.search(r"<SimFramework\b.*?</SimFramework>", text, re.S)
    if not m:
        raise ValueError(f"{path}: no <SimFramework> element")
    sim = ET.fromstring(m.group(0)).find("simulator")
    if sim is None:
        raise ValueError(f"{path}: no <simulator> element")
    p = {}
    extra = {}
    for el in sim:
        if el.tag in _PARAM_KEYS:
            p[_PARAM_KEYS[el.tag]] = _floats(el)[0]
        elif el.tag in _UNUSED_KEYS:
            extra[el.tag] = _floats(el)
    p.update(params)
    scene = scenes.make_scene(p)
    scene["extra"] = extra
    scene["surfaces"] = []
    scene["objects"] = []
    scene["model_order"] = []  # the reference adds simulated models in config order
    for el in sim:
        if el.tag not in ("object", "static_object"):
            continue
        attrs = list(el.attrib.items())
        if not attrs or attrs[0][0] != "name":
            raise ValueError(f"{path}: an <{el.tag}> without a leading name attribute")
        name = attrs[0][1]
        a = dict(attrs)
        cfg = dict(type=[1, 0.0, 0.0, 0.0], patchDim=[10], scale=[1, 1, 1], translation=[0, 0, 0],
                   rotation=[0, 0, 0, 0], density=[10.0])
        for ch in el:
            if ch.tag in cfg:
                cfg[ch.tag] = _floats(ch)
        model_type = int(cfg["type"][0])
        s0, s1, s2 = (cfg["type"] + [0.0, 0.0, 0.0])[1:4]
        d = os.path.join(root_dir, name)
        surf_v, surf_f = read_obj(os.path.join(d, name + ".obj"))
        nodes = tets = None
        if a.get("volumetric_mesh") == "tet":
            nodes, tets = read_tet(os.path.join(d, name + ".node"), os.path.join(d, name + ".ele"))
        sv, tv = transform(surf_v, nodes, _flag(a.get("uniform", "on")), _flag(a.get("origin", "on")),
                           cfg["scale"], cfg["translation"], cfg["rotation"])
        patch = int(cfg["patchDim"][0])
        fixed_path = os.path.join(d, "fixed_points.txt")
        fixed = read_fixed(fixed_path) if os.path.exists(fixed_path) else np.zeros(0, np.int32)
        scene["objects"].append(dict(name=name, kind=el.tag, model_type=model_type, dir=d))
        if el.tag == "static_object":
            if tets is not None:
                raise NotImplementedError(f"{name}: static volumetric objects are not supported")
            scenes.add_static(scene, sv.astype(np.float64), surf_f, patch_dim=patch)
        elif model_type == 1:
            if tets is None:
                raise ValueError(f"{name}: a volumetric type without volumetric_mesh=\"tet\"")
            scenes.add_tet(scene, tv.astype(np.float64), tets, strain=s0, vol=s1, pos=s2,
                           density=cfg["density"][0], patch_dim=patch, fixed=fixed)
            scene["model_order"].append(("tet", len(scene["tets"]) - 1))
            interp = os.path.join(d, "tetMeshInterpolation.txt")
            scene["surfaces"].append(dict(name=name, tet_model=len(scene["tets"]) - 1, v=sv, f=surf_f,
                                          interpolation=read_interpolation(interp)
                                          if os.path.exists(interp) else None))
        else:
            scenes.add_cloth(scene, sv.astype(np.float64), surf_f, strain=s0, bend=s1, pos=s2,
                             density=cfg["density"][0], patch_dim=patch, fixed=fixed)
            scene["model_order"].append(("cloth", len(scene["cloths"]) - 1))
    return scene


def surface_positions(scene, surface_index, tet_positions):
    """Surface vertices of a volumetric object from its tet node positions
    (BaseModel::updateSurfaceFromTetMesh): w0 x[t0] + w1 x[t1] + w2 x[t2] + w3 x[t3]
    in float32, summed in that order.  Without tetMeshInterpolation.txt the
    weights are the FP64 barycentric coordinates of the first tet (in tet
    order) that contains the rest vertex, or the nearest node's corner of its
    first tet (parity unpinned: the reference's own computation uses a float
    4x4 inverse)."""
    s = scene["surfaces"][surface_index]
    tets = scene["tets"][s["tet_model"]]["t"]
    if s["interpolation"] is None:
        s["interpolation"] = _interpolation(s["v"], scene["tets"][s["tet_model"]]["v"], tets)
    eid, w = s["interpolation"]
    x = np.asarray(tet_positions, F32)
    idx = tets[eid]
    out = w[:, :1] * x[idx[:, 0]]
    for k in range(1, 4):
        out = out + w[:, k:k + 1] * x[idx[:, k]]
    return out


def _interpolation(points, nodes, tets):
    """BaseModel::computeTetMeshInterpolation (Model/BaseModel.cpp:249-350):
    for each surface point the first tet (in tet order) whose barycentric
    coordinates are all >= 0 once those below 1e-6 in magnitude are set to 0
    (BaseTetMesh::isInsideTetElement, VolumetricMesh/BaseTetMesh.cpp:241-259,
    with the float build's MIN_VALUE), normalised to sum 1; otherwise the
    nearest tet node and the first tet that holds it, with that corner's weight
    1.  Coordinates come from an FP64 affine inverse per tet.  A uniform grid
    of tet boxes (padded by 1e-5 of the tet's extent, beyond what the 1e-6
    tolerance admits) lists each point's candidate tets in tet order."""
    pts = np.asarray(points, np.float64)
    nd = np.asarray(nodes, np.float64)
    tets = np.asarray(tets)
    X = nd[tets]  # (T, 4, 3)
    A = np.concatenate([X, np.ones((len(tets), 4, 1))], 2).transpose(0, 2, 1)  # columns: (x, 1) per corner
    inv = np.linalg.inv(A)
    ext = (X.max(1) - X.min(1)).max(1)
    lo = X.min(1) - (1e-5 * ext + 1e-12)[:, None]
    hi = X.max(1) + (1e-5 * ext + 1e-12)[:, None]
    g0 = np.minimum(lo.min(0), pts.min(0))
    h = max(float(np.median(ext)), 1e-12)
    c_lo = np.floor((lo - g0) / h).astype(np.int64)
    c_hi = np.floor((hi - g0) / h).astype(np.int64)
    cells = {}
    for t in range(len(tets)):
        for cx in range(c_lo[t, 0], c_hi[t, 0] + 1):
            for cy in range(c_lo[t, 1], c_hi[t, 1] + 1):
                for cz in range(c_lo[t, 2], c_hi[t, 2] + 1):
                    cells.setdefault((cx, cy, cz), []).append(t)
    pc = np.floor((pts - g0) / h).astype(np.int64)
    first_tet = {}
    for t, row in enumerate(tets):
        for v in row:
            first_tet.setdefault(int(v), t)
    eid = np.zeros(len(pts), np.int32)
    w = np.zeros((len(pts), 4))
    for i, p in enumerate(pts):
        cand = cells.get(tuple(pc[i]))
        j = -1
        if cand:
            cand = np.asarray(cand)
            inb = ((p >= lo[cand]) & (p <= hi[cand])).all(1)
            cand = cand[inb]
            if len(cand):
                b = inv[cand] @ np.append(p, 1.0)
                b[np.abs(b) < 1e-6] = 0.0
                ins = np.nonzero((b >= 0).all(1))[0]
                if len(ins):
                    j = int(cand[ins[0]])
                    eid[i], w[i] = j, b[ins[0]] / b[ins[0]].sum()
        if j < 0:
            nid = int(np.argmin(((nd - p) ** 2).sum(1)))
            j = first_tet[nid]
            eid[i] = j
            w[i] = (tets[j] == nid).astype(float)
    return eid, w.astype(F32)


def dat_cache_matches(sim, scene):
    """For each simulated cloth/volumetric model with cached boundary.dat /
    patchInfo.dat in its folder: whether the library's host precompute
    reproduces them (the reference reads these caches instead of recomputing
    when present, PdIpcModel.h:179-318)."""
    out = {}
    models = [o for o in scene["objects"] if o["kind"] == "object"]  # add order
    for mi, o in enumerate(models):
        for which, fname in ((0, "boundary.dat"), (1, "patchInfo.dat")):
            f = os.path.join(o["dir"], fname)
            if os.path.exists(f):
                out[(o["name"], fname)] = bool(np.array_equal(sim.model_dat(mi, which), np.fromfile(f, np.int32)))
    return out
