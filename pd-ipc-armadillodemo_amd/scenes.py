"""Seeded synthetic scenes for the PD-IPC step.

Every scene is a plain dict of numpy arrays and parameters, consumed unchanged
by the HIP library (pdipc.py) and by the CPU oracle, so both start from
bit-identical inputs.  Geometry follows the reference's conventions: gravity
acts along -y (PdIpcSimulator.cpp:249-252), so "height" is y and a cloth lies
in the x-z plane.

Randomness comes from numpy's PCG64 generator with an explicit seed.
"""

import numpy as np

# Reference defaults for the step constants that the reference hard-codes
# (PdIpcSimulator.cpp:1592,1597,1685,1776; PdIpcGpuFunc.cu:307-316,949,1318,
# 2656,2699,2497-2527; PdIpcSimulator.cpp:1494).
DEFAULT_PARAMS = dict(
    dt=0.01, gravity=-9.8, dhat=1e-3, kappa=1e4,
    lg_iters=10, cheb_iters=15, max_outer=30, residual_tol=1e-5,
    omega=2.0 / 3.0, rho=0.9992, sigma_lo=0.999, sigma_hi=1.001,
    ccd_filter=5e-3, toi_scale=0.98, ccd_iters=10, ccd_step=0.9, ccd_gap=0.1,
    static_mass=1e4, inflation=-1.0, skip_static_pairs=0, touch_policy=1,
)


def cloth_grid(n, side=0.5, height=0.2, jitter=2e-3, seed=0, center=(0.0, 0.0)):
    """n x n vertex square cloth in the x-z plane at y = height, with seeded
    uniform out-of-plane jitter in [-jitter, jitter]; 2 (n-1)^2 triangles with
    +y normals."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.linspace(-side / 2, side / 2, n)
    xs, zs = np.meshgrid(t, t, indexing="ij")
    v = np.zeros((n * n, 3))
    v[:, 0] = xs.reshape(-1) + center[0]
    v[:, 2] = zs.reshape(-1) + center[1]
    v[:, 1] = height + rng.uniform(-jitter, jitter, n * n)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (i * n + j).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    f = np.empty((2 * len(v00), 3), dtype=np.int32)
    f[0::2] = np.stack([v00, v01, v11], 1)
    f[1::2] = np.stack([v00, v11, v10], 1)
    return v, f


def _value_noise3(p, seed, octaves=3, base_freq=4.0):
    """Seeded 3-D value noise in [-1, 1]: random lattice values, smoothstep
    trilinear interpolation, octaves at doubling frequency and halving weight."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)

    def lattice(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791)
        return table[perm[h & 4095]]

    total = np.zeros(len(p))
    norm = 0.0
    amp = 1.0
    freq = base_freq
    for _ in range(octaves):
        q = p * freq + 1000.0
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        s = fr * fr * (3 - 2 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1 - s[:, 2]
                    acc += wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def icosphere(subdiv, radius=0.15, noise=0.015, seed=0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `subdiv` times (10*4^s+2 vertices, 20*4^s faces),
    projected to the sphere and displaced radially by seeded 3-octave value
    noise of the given amplitude.  Outward-oriented faces."""
    g = (1 + 5 ** 0.5) / 2
    verts = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g),
             (0, -1, -g), (0, 1, -g), (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    verts = [np.array(v, float) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    d = np.array(verts)
    r = radius + noise * _value_noise3(d, seed)
    v = d * r[:, None] + np.asarray(center)[None, :]
    return v, np.array(faces, dtype=np.int32)


def ground_plane(y=-0.2, half=1.0):
    """Two-triangle square at height y with +y normal."""
    v = np.array([[-half, y, -half], [-half, y, half], [half, y, half], [half, y, -half]], float)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return v, f


def tet_block(n, size=0.4, origin=(-0.2, -0.2, -0.2)):
    """n^3 cubes, 5 tets per cube with alternating parity (conforming), (n+1)^3
    vertices; every tet positively oriented."""
    m = n + 1
    t = np.arange(m)
    gx, gy, gz = np.meshgrid(t, t, t, indexing="ij")
    v = np.stack([gx, gy, gz], -1).reshape(-1, 3) * (size / n) + np.asarray(origin)[None, :]

    def vid(i, j, k):
        return (i * m + j) * m + k

    even = [(0, 1, 2, 4), (3, 1, 2, 7), (5, 1, 4, 7), (6, 2, 4, 7), (1, 2, 4, 7)]
    odd = [(1, 0, 3, 5), (2, 0, 3, 6), (4, 0, 5, 6), (7, 3, 5, 6), (0, 3, 5, 6)]
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)
    corners = np.stack([vid(ci + dx, cj + dy, ck + dz)
                        for dx, dy, dz in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0),
                                           (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]], 1)
    par = (ci + cj + ck) % 2
    tets = []
    for cell in range(len(ci)):
        pat = even if par[cell] == 0 else odd
        for tt in pat:
            tets.append(corners[cell, list(tt)])
    tets = np.array(tets, dtype=np.int32)
    a, b, c, d = (v[tets[:, k]] for k in range(4))
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    neg = det < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return v, tets


def make_scene(params=None):
    p = dict(DEFAULT_PARAMS)
    if params:
        p.update(params)
    return dict(params=p, cloths=[], tets=[], statics=[])


def add_cloth(scene, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=64, fixed=()):
    scene["cloths"].append(dict(v=np.ascontiguousarray(v, float), f=np.ascontiguousarray(f, np.int32),
                                s0=strain, s1=bend, s2=pos, density=density, patch_dim=patch_dim,
                                fixed=np.asarray(fixed, np.int32)))


def add_tet(scene, v, t, strain, vol=1e3, pos=1e12, density=1000.0, patch_dim=10, fixed=()):
    scene["tets"].append(dict(v=np.ascontiguousarray(v, float), t=np.ascontiguousarray(t, np.int32),
                              s0=strain, s1=vol, s2=pos, density=density, patch_dim=patch_dim,
                              fixed=np.asarray(fixed, np.int32)))


def add_static(scene, v, f, patch_dim=64):
    scene["statics"].append(dict(v=np.ascontiguousarray(v, float), f=np.ascontiguousarray(f, np.int32),
                                 patch_dim=patch_dim))


def cloth_over_sphere(n_cloth=64, subdiv=4, ground=False, seed=0, patch_dim=None, sphere_patch_dim=64,
                      height=0.2, kappa=1e3, params=None):
    """The benchmark family: an n x n cloth (0.5 m side, 2 mm seeded jitter)
    dropped from `height` onto a static noise-displaced icosphere (r = 0.15 m,
    1.5 cm noise), optionally above a ground plane.  PD strain weight 1e4,
    bending 1e-2, density 0.2 kg/m^2, dt 1/100, dHat 1e-3."""
    p = dict(kappa=kappa)
    if params:
        p.update(params)
    s = make_scene(p)
    v, f = cloth_grid(n_cloth, side=0.5, height=height, jitter=2e-3, seed=seed)
    add_cloth(s, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2,
              patch_dim=patch_dim if patch_dim else n_cloth)
    sv, sf = icosphere(subdiv, radius=0.15, noise=0.015, seed=seed)
    add_static(s, sv, sf, patch_dim=sphere_patch_dim)
    if ground:
        gv, gf = ground_plane(y=-0.2, half=1.0)
        add_static(s, gv, gf, patch_dim=2)
    return s


def config0(seed=0):
    """64x64 cloth over a subdivision-4 noisy icosphere (7,938 + 5,120 triangles)."""
    return cloth_over_sphere(64, 4, ground=False, seed=seed, patch_dim=64, sphere_patch_dim=64)


def config1(seed=0):
    """256x256 cloth over a subdivision-6 noisy icosphere plus a ground plane
    (130,050 + 81,920 + 2 triangles)."""
    return cloth_over_sphere(256, 6, ground=True, seed=seed, patch_dim=256, sphere_patch_dim=256)


def config2(seed=0, cells=40, cloth_n=128):
    """40^3-cell tet block (5 tets/cell, E = 1e5 Pa, nu = 0.4, rho = 1000) on the
    ground with a 128x128 cloth over it; dt = 1/200, 15 local/global iterations.
    The PD strain weight is 2 mu = E / (1 + nu)."""
    E, nu = 1e5, 0.4
    s = make_scene(dict(dt=1.0 / 200.0, lg_iters=15, kappa=1e3))
    cv, cf = cloth_grid(cloth_n, side=0.6, height=0.23, jitter=2e-3, seed=seed)
    add_cloth(s, cv, cf, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=cloth_n)
    tv, tt = tet_block(cells, size=0.4, origin=(-0.2, -0.19, -0.2))
    add_tet(s, tv, tt, strain=E / (1 + nu), vol=1e3, pos=1e12, density=1000.0, patch_dim=max(cells, 2))
    gv, gf = ground_plane(y=-0.2, half=1.0)
    add_static(s, gv, gf, patch_dim=2)
    return s


def config3(seed=0, n_cloth=128, steps=50):
    """The differentiable run: a 128 x 128 cloth (32,258 triangles) over the
    subdivision-4 icosphere obstacle, FP64, `steps` steps.  Returns
    (scene, v0, target): per-triangle strain stiffness log-uniform in
    [1e3, 1e5] (an edge's spring weight is the mean of its triangles), initial
    velocity N(0, 0.05^2) per component, and a seeded target drape (the rest
    grid lowered by 5 cm plus N(0, 1 cm^2) noise) for the loss
    sum |x_T - target|^2."""
    s = cloth_over_sphere(n_cloth, 4, ground=False, seed=seed, patch_dim=64, sphere_patch_dim=64)
    rng = np.random.default_rng(1_000_003 + seed)
    c = s["cloths"][0]
    c["tri_k"] = np.exp(rng.uniform(np.log(1e3), np.log(1e5), len(c["f"])))
    v0 = rng.normal(0.0, 0.05, c["v"].shape)
    target = c["v"] + np.array([0.0, -0.05, 0.0]) + rng.normal(0.0, 0.01, c["v"].shape)
    s["steps"] = steps
    return s, v0, target


def config4(seed=0, n=1024, layers=3, spacing=5e-3, side=2.0):
    """The scale test: `layers` stacked n x n cloths (3 x 1024^2 = 3,145,728
    vertices, 6,279,174 triangles), `spacing` apart, each with a seeded
    in-plane shear x += s z, |s| <= 0.1, over the subdivision-6 sphere and a
    3 m ground plane.  The side is 2 m so the grid spacing (1.96 mm) is that of
    the 256^2 cloth on 0.5 m (an edge shorter than dHat = 1 mm would put
    barrier pairs inside each sheet); the out-of-plane jitter is 1 mm, so
    adjacent layers start at least 3 mm apart."""
    s = make_scene(dict(kappa=1e3))
    rng = np.random.Generator(np.random.PCG64(7_000_003 + seed))
    for k in range(layers):
        v, f = cloth_grid(n, side=side, height=0.2 + k * spacing, jitter=1e-3, seed=seed + 101 * k)
        shear = rng.uniform(-0.1, 0.1)
        v[:, 0] += shear * v[:, 2]
        add_cloth(s, v, f, strain=1e4, bend=1e-2, pos=1e12, density=0.2, patch_dim=n)
    sv, sf = icosphere(6, radius=0.15, noise=0.015, seed=seed)
    add_static(s, sv, sf, patch_dim=256)
    gv, gf = ground_plane(y=-0.2, half=1.5)
    add_static(s, gv, gf, patch_dim=2)
    return s


def scene_sizes(scene):
    nv = sum(len(c["v"]) for c in scene["cloths"]) + sum(len(t["v"]) for t in scene["tets"])
    nf = sum(len(c["f"]) for c in scene["cloths"]) + sum(len(t["f"]) for t in scene["statics"])
    nt = sum(len(t["t"]) for t in scene["tets"])
    ns = sum(len(t["v"]) for t in scene["statics"])
    return dict(sim_points=nv, static_points=ns, tets=nt, surface_triangles=nf)
