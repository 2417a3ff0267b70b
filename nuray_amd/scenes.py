"""Seeded synthetic scenes for the ptnee workloads.

Every generator uses only IEEE-exact numpy arithmetic (+, -, *, /, sqrt) on float64 and seeded
PCG64 streams, then rounds to float32, so the triangle arrays are bit-identical on every host.

A scene is a dict of numpy arrays in the C-ABI layout (see include/nuray.h):
  p    (N, 9) float32   triangle vertices p0, p1, p2 in loader order
  n    (N, 9) float32 or None   vertex normals (None -> flat normals, triangle.cpp:14-17)
  uv   (N, 6) float32 or None   texture coordinates
  mat  (N,)   int32     material row per triangle
  mats (M, 8) float32   {type, r, g, b, alpha, ior, texture, 0}
  camera dict(pos, euler, fov, aspect), image size, envmap (optional).
"""
from __future__ import annotations

import numpy as np

LAMBERT, LIGHT, GGXREFL, GGX = 0, 1, 2, 3

DEG = np.pi / 180.0


def mat_row(kind, rgb, alpha=0.5, ior=1.0, tex=-1):
    return [float(kind), float(rgb[0]), float(rgb[1]), float(rgb[2]), float(alpha), float(ior), float(tex), 0.0]


def _quad(p0, p1, p2, p3, normal):
    """Two triangles (p0,p1,p2), (p2,p3,p0) -- the loader's quad split (loader.cpp:336-354) --
    wound so that (p1-p0) x (p2-p0) points along `normal`."""
    p0, p1, p2, p3 = (np.asarray(v, np.float64) for v in (p0, p1, p2, p3))
    if np.dot(np.cross(p1 - p0, p2 - p0), normal) < 0:
        p1, p3 = p3, p1
    return [np.concatenate([p0, p1, p2]), np.concatenate([p2, p3, p0])]


def _rot_y(v, deg):
    c, s = _cos_sin(deg)
    x, y, z = v
    return np.array([c * x + s * z, y, -s * x + c * z])


def _cos_sin(deg):
    # cos/sin from a short Taylor series evaluated with + and * only (bit-reproducible on any host)
    a = deg * (3.141592653589793 / 180.0)
    a2 = a * a
    c, s, term_c, term_s = 1.0, a, 1.0, a
    for k in range(1, 12):
        term_c = -term_c * a2 / ((2 * k - 1) * (2 * k))
        term_s = -term_s * a2 / ((2 * k) * (2 * k + 1))
        c += term_c
        s += term_s
    return c, s


def _box(center, size, rot_deg, ntri_mats=None):
    """Closed axis-aligned box (12 triangles, faces outward) rotated about +y through its centre."""
    cx, cy, cz = center
    hx, hy, hz = (s / 2.0 for s in size)
    faces = []
    # (normal, four corners in local coordinates)
    for axis in range(3):
        for sign in (-1.0, 1.0):
            nrm = np.zeros(3)
            nrm[axis] = sign
            u = (axis + 1) % 3
            v = (axis + 2) % 3
            h = np.array([hx, hy, hz])
            corners = []
            for du, dv in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                q = np.zeros(3)
                q[axis] = sign * h[axis]
                q[u] = du * h[u]
                q[v] = dv * h[v]
                corners.append(q)
            world = [_rot_y(q, rot_deg) + np.array([cx, cy, cz]) for q in corners]
            faces.append((axis, sign, _quad(*world, _rot_y(nrm, rot_deg))))
    return faces


def box_room(patches: bool = False, seed: int = 5):
    """The closed box room: 200-unit room, five Lambert walls, a 50x50 ceiling light with emission
    (17, 12, 4), a 60^3 box rotated 15 deg and a 60x120x60 box rotated -18 deg (36 triangles).
    Camera (0,100,350), looking down -z, vertical fov 40, 320x240.

    With patches=True the walls and boxes are split into 64 GGX-reflect material regions (the
    differentiable-rendering scene); their hidden albedo / roughness come from `seed`."""
    white, red, green = (0.725, 0.71, 0.68), (0.63, 0.065, 0.05), (0.14, 0.45, 0.09)
    tris, mat = [], []
    mats = []

    def add(tlist, m):
        for t in tlist:
            tris.append(t)
            mat.append(m)

    R, H = 100.0, 200.0
    walls = [  # (corners, inward normal, colour, grid for patch mode)
        (((-R, 0, R), (R, 0, R), (R, 0, -R), (-R, 0, -R)), (0, 1, 0), white, (4, 4)),      # floor
        (((-R, H, R), (R, H, R), (R, H, -R), (-R, H, -R)), (0, -1, 0), white, (4, 2)),     # ceiling
        (((-R, 0, -R), (R, 0, -R), (R, H, -R), (-R, H, -R)), (0, 0, 1), white, (4, 4)),     # back
        (((-R, 0, R), (-R, 0, -R), (-R, H, -R), (-R, H, R)), (1, 0, 0), red, (2, 4)),       # left
        (((R, 0, -R), (R, 0, R), (R, H, R), (R, H, -R)), (-1, 0, 0), green, (2, 4)),       # right
    ]
    if not patches:
        mats = [mat_row(LAMBERT, white), mat_row(LAMBERT, red), mat_row(LAMBERT, green),
                mat_row(LIGHT, (17.0, 12.0, 4.0))]
        colour_id = {white: 0, red: 1, green: 2}
        for corners, nrm, col, _ in walls:
            add(_quad(*corners, np.array(nrm, float)), colour_id[col])
        light_id, box_mat = 3, lambda face_axis, sign: 0
    else:
        rng = np.random.Generator(np.random.PCG64(seed))
        n_region = 64
        albedo = 0.15 + 0.8 * rng.random((n_region, 3))
        rough = 0.15 + 0.7 * rng.random(n_region)
        mats = [mat_row(GGXREFL, albedo[i], rough[i]) for i in range(n_region)]
        mats.append(mat_row(LIGHT, (17.0, 12.0, 4.0)))
        light_id = n_region
        region = 0
        for corners, nrm, col, (gu, gv) in walls:
            c = [np.asarray(v, np.float64) for v in corners]
            for i in range(gu):
                for j in range(gv):
                    def P(a, b):
                        return c[0] + (c[1] - c[0]) * a + (c[3] - c[0]) * b
                    q = (P(i / gu, j / gv), P((i + 1) / gu, j / gv), P((i + 1) / gu, (j + 1) / gv),
                         P(i / gu, (j + 1) / gv))
                    add(_quad(*q, np.array(nrm, float)), region)
                    region += 1
        assert region == 56
        box_regions = {"short": 56, "tall": 60}

        def box_mat_fn(which):
            base = box_regions[which]

            def f(axis, sign):
                if axis == 1:
                    return base + (0 if sign > 0 else 3)
                return base + (1 if axis == 0 else 2)
            return f
        box_mat = box_mat_fn
    # ceiling light: 50 x 50 at y = 199, facing down
    add(_quad((-25, 199, -25), (25, 199, -25), (25, 199, 25), (-25, 199, 25), np.array([0, -1.0, 0])), light_id)
    for which, center, size, rot in (("short", (40.0, 30.0, 30.0), (60.0, 60.0, 60.0), 15.0),
                                      ("tall", (-35.0, 60.0, -35.0), (60.0, 120.0, 60.0), -18.0)):
        for axis, sign, tl in _box(center, size, rot):
            add(tl, box_mat(which)(axis, sign) if patches else 0)
    p = np.asarray(tris, np.float64).astype(np.float32)
    scene = dict(p=p, n=None, uv=None, mat=np.asarray(mat, np.int32), mats=np.asarray(mats, np.float32),
                 camera=dict(pos=(0.0, 100.0, 350.0), euler=(0.0, 0.0, 0.0), fov=40.0, aspect=4.0 / 3.0),
                 width=320, height=240, envmap=None, name="box_room" + ("_64mat" if patches else ""))
    return scene


def _icosphere80():
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = v / np.sqrt((v * v).sum(1, keepdims=True))
    verts = list(v)
    cache = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key not in cache:
            m = (verts[a] + verts[b]) / 2.0
            verts.append(m / np.sqrt((m * m).sum()))
            cache[key] = len(verts) - 1
        return cache[key]
    faces = []
    for a, b, c in f:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    return np.asarray(verts), np.asarray(faces)


def _value_noise(x, z, seed, octaves=5, base=8):
    """fBm of bilinear-smoothstep value noise on a square lattice; x, z in [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros_like(x)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        res = base << o
        lat = rng.random((res + 1, res + 1)) * 2.0 - 1.0
        gx, gz = x * res, z * res
        ix = np.minimum(np.floor(gx).astype(np.int64), res - 1)
        iz = np.minimum(np.floor(gz).astype(np.int64), res - 1)
        fx, fz = gx - ix, gz - iz
        sx = fx * fx * (3.0 - 2.0 * fx)
        sz = fz * fz * (3.0 - 2.0 * fz)
        a = lat[ix, iz] * (1 - sx) + lat[ix + 1, iz] * sx
        b = lat[ix, iz + 1] * (1 - sx) + lat[ix + 1, iz + 1] * sx
        out = out + amp * (a * (1 - sz) + b * sz)
        total += amp
        amp *= 0.5
    return out / total


def _sky(w=1024, h=512):
    """Procedural sky, RGB float32, polynomial in the elevation (row index) and azimuth."""
    j = (np.arange(h, dtype=np.float64) + 0.5) / h          # texture row -> 1 - v
    i = (np.arange(w, dtype=np.float64) + 0.5) / w          # column -> u
    v = 1.0 - j                                             # v = 1 - acos(y)/pi: v > 0.5 above the horizon
    elev = np.clip((v - 0.5) * 2.0, -1.0, 1.0)[:, None]
    up = np.clip(elev, 0.0, 1.0)
    zenith = np.array([0.25, 0.45, 0.9])
    horizon = np.array([0.85, 0.8, 0.7])
    ground = np.array([0.25, 0.22, 0.2])
    sky = horizon + (zenith - horizon) * (up * (2.0 - up))[..., None]
    sky = np.where((elev < 0)[..., None], ground + (horizon - ground) * (1.0 + elev[..., None]) ** 4, sky)
    sky = np.broadcast_to(sky, (h, w, 3)).copy()
    # a soft "sun" lobe: quartic bump around (u, v) = (0.3, 0.78)
    du = (i[None, :] - 0.3) * 4.0
    dv = (v[:, None] - 0.78) * 8.0
    r2 = du * du + dv * dv
    bump = np.clip(1.0 - r2, 0.0, 1.0) ** 4
    sky = sky + (np.array([30.0, 27.0, 22.0]) * bump[..., None])
    return sky.astype(np.float32)


def terrain_spheres(n_cells: int = 512, n_spheres: int = 2000, width=1280, height=720):
    """The procedural heightfield + icosphere scene: n_cells x n_cells cell heightfield over
    x, z in [0, 1000] with y = 20 * valueNoise(seed=1) (Lambert, smooth vertex normals),
    `n_spheres` 80-triangle icospheres (r ~ U[1,8], centre ~ U([0,1000]x[0,200]x[0,1000]), seed 2),
    each a GGX-reflect material with roughness ~ U[0.05, 0.8] and albedo ~ U[0.1, 0.9]^3 (seed 3),
    four 150x150 emissive quads at y = 400, and a 1024x512 procedural sky envmap."""
    m = n_cells
    g = np.arange(m + 1, dtype=np.float64) / m
    X, Z = np.meshgrid(g, g, indexing="ij")
    Y = 20.0 * _value_noise(X, Z, seed=1)
    P = np.stack([X * 1000.0, Y, Z * 1000.0], -1)
    # smooth vertex normals by central differences
    dx = np.zeros_like(Y)
    dz = np.zeros_like(Y)
    dx[1:-1, :] = (Y[2:, :] - Y[:-2, :]) / (2 * 1000.0 / m)
    dx[0, :] = (Y[1, :] - Y[0, :]) / (1000.0 / m)
    dx[-1, :] = (Y[-1, :] - Y[-2, :]) / (1000.0 / m)
    dz[:, 1:-1] = (Y[:, 2:] - Y[:, :-2]) / (2 * 1000.0 / m)
    dz[:, 0] = (Y[:, 1] - Y[:, 0]) / (1000.0 / m)
    dz[:, -1] = (Y[:, -1] - Y[:, -2]) / (1000.0 / m)
    Nn = np.stack([-dx, np.ones_like(Y), -dz], -1)
    Nn = Nn / np.sqrt((Nn * Nn).sum(-1, keepdims=True))
    a = P[:-1, :-1].reshape(-1, 3); b = P[1:, :-1].reshape(-1, 3)
    c = P[1:, 1:].reshape(-1, 3); d = P[:-1, 1:].reshape(-1, 3)
    na = Nn[:-1, :-1].reshape(-1, 3); nb = Nn[1:, :-1].reshape(-1, 3)
    nc = Nn[1:, 1:].reshape(-1, 3); nd = Nn[:-1, 1:].reshape(-1, 3)
    # quad (a, d, c, b) is wound with +y facing normal: (d-a) x (c-a) has +y
    t1 = np.concatenate([a, d, c], 1); t2 = np.concatenate([c, b, a], 1)
    n1 = np.concatenate([na, nd, nc], 1); n2 = np.concatenate([nc, nb, na], 1)
    tp = np.stack([t1, t2], 1).reshape(-1, 9)
    tn = np.stack([n1, n2], 1).reshape(-1, 9)
    mats = [mat_row(LAMBERT, (0.55, 0.5, 0.42))]
    tmat = np.zeros(len(tp), np.int32)

    rng2 = np.random.Generator(np.random.PCG64(2))
    rng3 = np.random.Generator(np.random.PCG64(3))
    radius = 1.0 + 7.0 * rng2.random(n_spheres)
    centre = rng2.random((n_spheres, 3)) * np.array([1000.0, 200.0, 1000.0])
    rough = 0.05 + 0.75 * rng3.random(n_spheres)
    alb = 0.1 + 0.8 * rng3.random((n_spheres, 3))
    sv, sf = _icosphere80()
    sp_p = (sv[sf].reshape(1, 80, 9) * np.repeat(radius, 9).reshape(n_spheres, 1, 9)
            + np.tile(centre, (1, 3)).reshape(n_spheres, 1, 9)).reshape(-1, 9)
    sp_n = np.tile(sv[sf].reshape(80, 9), (n_spheres, 1))
    base = len(mats)
    for k in range(n_spheres):
        mats.append(mat_row(GGXREFL, alb[k], rough[k]))
    sp_mat = (base + np.repeat(np.arange(n_spheres), 80)).astype(np.int32)

    light_id = len(mats)
    mats.append(mat_row(LIGHT, (6.0, 5.8, 5.2)))
    lp = []
    for cx, cz in ((250.0, 250.0), (750.0, 250.0), (250.0, 750.0), (750.0, 750.0)):
        lp += _quad((cx - 75, 400, cz - 75), (cx + 75, 400, cz - 75), (cx + 75, 400, cz + 75),
                    (cx - 75, 400, cz + 75), np.array([0, -1.0, 0]))
    lp = np.asarray(lp)
    ln = np.tile(np.array([0, -1.0, 0] * 3), (len(lp), 1))

    p = np.concatenate([tp, sp_p, lp]).astype(np.float32)
    n = np.concatenate([tn, sp_n, ln]).astype(np.float32)
    mat = np.concatenate([tmat, sp_mat, np.full(len(lp), light_id, np.int32)])
    return dict(p=p, n=n, uv=None, mat=mat, mats=np.asarray(mats, np.float32),
                camera=dict(pos=(500.0, 300.0, 1150.0), euler=(0.0, -28.0, 0.0), fov=45.0,
                            aspect=width / height),
                width=width, height=height, envmap=dict(rgb=_sky(), gain=1.0),
                name=f"terrain{n_cells}_spheres{n_spheres}")


def camera_array(cam):
    """{pos, euler, fov, aspect} -> float32[8] as the C ABI takes it."""
    return np.array(list(cam["pos"]) + list(cam["euler"]) + [cam["fov"], cam["aspect"]], np.float32)


def theta_of(scene):
    """Material parameter rows {r, g, b, alpha} of a scene (emitters included, ignored on set)."""
    m = scene["mats"]
    return np.ascontiguousarray(np.stack([m[:, 1], m[:, 2], m[:, 3], m[:, 4]], 1), np.float32)


def adjoint_problem(seed_hidden=5, seed_start=6, width=256, height=256):
    """The inverse-rendering problem: the 64-region GGX-reflect box room with hidden parameters
    theta* (seed 5), and a start theta0 = theta* * (1 + U[-0.3, 0.3]) per parameter (seed 6).
    Returns (scene at theta*, theta*, theta0); rows of the emitter are left unchanged."""
    sc = box_room(patches=True, seed=seed_hidden)
    sc["width"], sc["height"] = width, height
    sc["camera"] = dict(sc["camera"], aspect=width / height)
    theta_star = theta_of(sc)
    rng = np.random.Generator(np.random.PCG64(seed_start))
    pert = 1.0 + rng.uniform(-0.3, 0.3, theta_star.shape)
    is_light = sc["mats"][:, 0] == LIGHT
    theta0 = np.where(is_light[:, None], theta_star, theta_star * pert).astype(np.float32)
    theta0[:, :3] = np.clip(theta0[:, :3], 0.0, None)
    return sc, theta_star, theta0


def _checker_rgba(w, h, seed):
    """RGBA8 checkerboard with a colour gradient (seeded tints)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tint = 0.3 + 0.7 * rng.random((2, 3))
    y, x = np.mgrid[0:h, 0:w]
    chk = ((x // 8 + y // 8) % 2).astype(np.float64)
    grad = (x / max(w - 1, 1))[..., None]
    rgb = (chk[..., None] * tint[0] + (1 - chk[..., None]) * tint[1]) * (0.5 + 0.5 * grad)
    out = np.zeros((h, w, 4), np.uint8)
    out[..., :3] = np.floor(rgb * 255.0 + 0.5).astype(np.uint8)
    out[..., 3] = 255
    return out


def material_zoo(width=160, height=120):
    """The box room with every material feature of the path: a textured Lambert floor (uv 0..4,
    wrapping), a textured GGX-reflect back wall, Lambert side walls, two translucent GGX dielectric
    icospheres with smooth normals (alpha 0.2 / ior 1.5 and alpha 0.5 / ior 1.3), a GGX-reflect tall
    box, the ceiling light, and a 64x32 sky through the open front."""
    base = box_room()
    tris, n9, uv6, mat = [], [], [], []
    white, red, green = (0.725, 0.71, 0.68), (0.63, 0.065, 0.05), (0.14, 0.45, 0.09)
    mats = [mat_row(LAMBERT, white, tex=0),            # 0 floor, textured
            mat_row(GGXREFL, (1, 1, 1), 0.3, tex=1),   # 1 back wall, textured F0
            mat_row(LAMBERT, red), mat_row(LAMBERT, green), mat_row(LAMBERT, white),  # 2, 3, 4
            mat_row(LIGHT, (17.0, 12.0, 4.0)),         # 5
            mat_row(GGX, (0.95, 0.95, 0.95), 0.2, ior=1.5),   # 6 glass-like sphere
            mat_row(GGX, (0.9, 0.8, 0.7), 0.5, ior=1.3),      # 7 rough dielectric sphere
            mat_row(GGXREFL, (0.8, 0.7, 0.5), 0.15)]          # 8 tall box
    R, H = 100.0, 200.0

    def add_quad(corners, nrm, m, uvq=None):
        tl = _quad(*corners, np.array(nrm, float))
        uvq = uvq if uvq is not None else [(0, 0), (0, 0), (0, 0), (0, 0)]
        # _quad may swap p1 / p3 to fix the winding: recover the corner order from the points
        cs = [np.asarray(c, np.float64) for c in corners]
        for t in tl:
            pts = t.reshape(3, 3)
            idx = [int(np.argmin([np.abs(pp - c).sum() for c in cs])) for pp in pts]
            tris.append(t)
            n9.append(np.zeros(9))
            uv6.append(np.concatenate([np.asarray(uvq[i], np.float64) for i in idx]))
            mat.append(m)
    add_quad(((-R, 0, R), (R, 0, R), (R, 0, -R), (-R, 0, -R)), (0, 1, 0), 0, [(0, 0), (4, 0), (4, 4), (0, 4)])
    add_quad(((-R, 0, -R), (R, 0, -R), (R, H, -R), (-R, H, -R)), (0, 0, 1), 1, [(0, 0), (1, 0), (1, 1), (0, 1)])
    add_quad(((-R, 0, R), (-R, 0, -R), (-R, H, -R), (-R, H, R)), (1, 0, 0), 2)
    add_quad(((R, 0, -R), (R, 0, R), (R, H, R), (R, H, -R)), (-1, 0, 0), 3)
    add_quad(((-R, H, R), (R, H, R), (R, H, -R), (-R, H, -R)), (0, -1, 0), 4)
    add_quad(((-25, 199, -25), (25, 199, -25), (25, 199, 25), (-25, 199, 25)), (0, -1, 0), 5)
    for axis, sign, tl in _box((-35.0, 60.0, -35.0), (60.0, 120.0, 60.0), -18.0):
        for t in tl:
            tris.append(t); n9.append(np.zeros(9)); uv6.append(np.zeros(6)); mat.append(8)
    sv, sf = _icosphere80()
    for (c, r, m) in (((40.0, 30.0, 30.0), 30.0, 6), ((-50.0, 25.0, 60.0), 25.0, 7)):
        for f in sf:
            pts = sv[f] * r + np.asarray(c)
            tris.append(pts.reshape(-1)); n9.append(sv[f].reshape(-1)); uv6.append(np.zeros(6)); mat.append(m)
    sky = _sky(64, 32) * 0.5
    return dict(p=np.asarray(tris).astype(np.float32), n=np.asarray(n9).astype(np.float32),
                uv=np.asarray(uv6).astype(np.float32), mat=np.asarray(mat, np.int32),
                mats=np.asarray(mats, np.float32), textures=[_checker_rgba(64, 64, 1), _checker_rgba(32, 32, 2)],
                camera=dict(base["camera"], aspect=width / height), width=width, height=height,
                envmap=dict(rgb=sky, gain=1.0), name="material_zoo")


def textured_room(width=64, height=48, floor_tex=32, wall_tex=16, seed=4):
    """The box room with linear float albedo textures for the inverse-texture problem: the floor is
    Lambert with a floor_tex^2 texture (uv 0..2, wrapping), the back wall GGX-reflect (alpha 0.3)
    with a wall_tex^2 F0 texture; other walls and both boxes Lambert; the ceiling light."""
    sc = material_zoo(width, height)
    rng = np.random.Generator(np.random.PCG64(seed))
    keep = np.isin(sc["mat"], [0, 1, 2, 3, 4, 5, 8])       # drop the dielectric spheres
    sc["p"], sc["n"], sc["uv"], sc["mat"] = sc["p"][keep], sc["n"][keep], sc["uv"][keep], sc["mat"][keep]
    sc["uv"][sc["mat"] == 0] *= 0.5
    mats = sc["mats"].copy()
    mats[8] = mat_row(LAMBERT, (0.7, 0.6, 0.5))
    sc["mats"] = mats
    sc["textures"] = [(0.2 + 0.6 * rng.random((floor_tex, floor_tex, 3))).astype(np.float32),
                      (0.3 + 0.6 * rng.random((wall_tex, wall_tex, 3))).astype(np.float32)]
    sc["envmap"] = None
    sc["name"] = "textured_room"
    return sc


def _tex_rgba8(size, k):
    """Procedural RGBA8 albedo texture `k`: tinted 64-texel checker with a diagonal ramp."""
    rng = np.random.Generator(np.random.PCG64(100 + k))
    tint = (60 + 190 * rng.random((2, 3))).astype(np.float32)
    idx = np.arange(size, dtype=np.int32)
    chk = ((idx[:, None] >> 6) + (idx[None, :] >> 6)) & 1
    ramp = ((idx[:, None] + idx[None, :]).astype(np.float32) * (0.5 / (2 * size)) + 0.5)
    out = np.empty((size, size, 4), np.uint8)
    for c in range(3):
        v = np.where(chk == 1, tint[0, c], tint[1, c]) * ramp
        out[..., c] = np.clip(v, 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return out


def large_textured(n_cells=2048, n_spheres=10000, tex_size=4096, n_tex=16, adj_tex=2048, width=3840, height=2160):
    """The large textured scene: an n_cells^2 heightfield over x, z in [0, 2000] (amplitude 30,
    value noise seed 1) split into 4 x 4 Lambert regions, each with its own tex_size^2 RGBA8 albedo
    texture (uv = position / 250, wrapping) except region 5, which carries the adj_tex^2 linear
    float texture that the adjoint optimises; n_spheres 80-triangle icospheres (r ~ U[2, 16],
    centres in [0,2000] x [0,300] x [0,2000], seed 2), 10 % translucent GGX dielectric (ior 1.5,
    alpha ~ U[0.05, 0.5]) and the rest GGX-reflect (alpha ~ U[0.05, 0.8], albedo ~ U[0.1, 0.9]^3,
    seed 3); four 300 x 300 emissive quads at y = 600 and the 1024 x 512 sky."""
    m = n_cells
    g = np.arange(m + 1, dtype=np.float64) / m
    X, Z = np.meshgrid(g, g, indexing="ij")
    Y = 30.0 * _value_noise(X, Z, seed=1)
    ext = 2000.0
    P = np.stack([X * ext, Y, Z * ext], -1)
    h = ext / m
    dx = np.gradient(Y, h, axis=0)
    dz = np.gradient(Y, h, axis=1)
    Nn = np.stack([-dx, np.ones_like(Y), -dz], -1)
    Nn /= np.sqrt((Nn * Nn).sum(-1, keepdims=True))
    UV = np.stack([X * ext / 250.0, Z * ext / 250.0], -1)
    a, b, c, d = (P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:])
    na, nb, nc, nd = (Nn[:-1, :-1], Nn[1:, :-1], Nn[1:, 1:], Nn[:-1, 1:])
    ua, ub, uc, ud = (UV[:-1, :-1], UV[1:, :-1], UV[1:, 1:], UV[:-1, 1:])
    t1 = np.concatenate([a, d, c], -1).reshape(-1, 9)
    t2 = np.concatenate([c, b, a], -1).reshape(-1, 9)
    tp = np.stack([t1, t2], 1).reshape(-1, 9)
    tn = np.stack([np.concatenate([na, nd, nc], -1).reshape(-1, 9),
                   np.concatenate([nc, nb, na], -1).reshape(-1, 9)], 1).reshape(-1, 9)
    tuv = np.stack([np.concatenate([ua, ud, uc], -1).reshape(-1, 6),
                    np.concatenate([uc, ub, ua], -1).reshape(-1, 6)], 1).reshape(-1, 6)
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    region = ((ii * 4 // m) * 4 + (jj * 4 // m)).reshape(-1)
    tmat = np.repeat(region, 2).astype(np.int32)
    mats = [mat_row(LAMBERT, (1, 1, 1), tex=(n_tex if r == 5 else r)) for r in range(16)]
    rng2 = np.random.Generator(np.random.PCG64(2))
    rng3 = np.random.Generator(np.random.PCG64(3))
    radius = 2.0 + 14.0 * rng2.random(n_spheres)
    centre = rng2.random((n_spheres, 3)) * np.array([ext, 300.0, ext])
    glass = rng3.random(n_spheres) < 0.1
    rough = np.where(glass, 0.05 + 0.45 * rng3.random(n_spheres), 0.05 + 0.75 * rng3.random(n_spheres))
    alb = np.where(glass[:, None], 0.8 + 0.2 * rng3.random((n_spheres, 3)), 0.1 + 0.8 * rng3.random((n_spheres, 3)))
    base = len(mats)
    for k in range(n_spheres):
        mats.append(mat_row(GGX, alb[k], rough[k], ior=1.5) if glass[k] else mat_row(GGXREFL, alb[k], rough[k]))
    sv, sf = _icosphere80()
    sp_p = (sv[sf].reshape(1, 80, 9) * np.repeat(radius, 9).reshape(n_spheres, 1, 9)
            + np.tile(centre, (1, 3)).reshape(n_spheres, 1, 9)).reshape(-1, 9)
    sp_n = np.tile(sv[sf].reshape(80, 9), (n_spheres, 1))
    sp_mat = (base + np.repeat(np.arange(n_spheres), 80)).astype(np.int32)
    light_id = len(mats)
    mats.append(mat_row(LIGHT, (8.0, 7.6, 7.0)))
    lp = []
    for cx, cz in ((500.0, 500.0), (1500.0, 500.0), (500.0, 1500.0), (1500.0, 1500.0)):
        lp += _quad((cx - 150, 600, cz - 150), (cx + 150, 600, cz - 150), (cx + 150, 600, cz + 150),
                    (cx - 150, 600, cz + 150), np.array([0, -1.0, 0]))
    lp = np.asarray(lp)
    ln = np.tile(np.array([0, -1.0, 0] * 3), (len(lp), 1))
    p = np.concatenate([tp, sp_p, lp]).astype(np.float32)
    n = np.concatenate([tn, sp_n, ln]).astype(np.float32)
    uv = np.concatenate([tuv, np.zeros((len(sp_p) + len(lp), 6))]).astype(np.float32)
    mat = np.concatenate([tmat, sp_mat, np.full(len(lp), light_id, np.int32)])
    rng4 = np.random.Generator(np.random.PCG64(4))
    textures = [_tex_rgba8(tex_size, k) for k in range(n_tex)]
    textures.append((0.2 + 0.6 * rng4.random((adj_tex, adj_tex, 3))).astype(np.float32))
    return dict(p=p, n=n, uv=uv, mat=mat, mats=np.asarray(mats, np.float32), textures=textures,
                camera=dict(pos=(1000.0, 500.0, 2300.0), euler=(0.0, -25.0, 0.0), fov=45.0, aspect=width / height),
                width=width, height=height, envmap=dict(rgb=_sky(), gain=1.0),
                name=f"large{n_cells}_spheres{n_spheres}_tex{n_tex}x{tex_size}")


BLINNPHONG, GLASS = 4, 5


def material_zoo_ext(width=160, height=120):
    """The box room with the remaining reference materials: a Blinn-Phong back wall (Kd from the
    floor's texture, Ks 0.5, Ns 50), a glass sphere (F0 0.04, ior 1.5), a Blinn-Phong sphere with a
    textured Ks and Ns 200, a GGX-reflect tall box, the textured Lambert floor, Lambert side walls,
    the ceiling light and the sky."""
    base = material_zoo(width, height)
    keep = ~np.isin(base["mat"], [6, 7])                       # drop the GGX dielectric spheres
    p, n, uv, mat = base["p"][keep], base["n"][keep], base["uv"][keep], base["mat"][keep].copy()
    mats = base["mats"].copy()
    mats[1] = mat_row(BLINNPHONG, (0.6, 0.5, 0.4), 50.0, tex=0)   # back wall: Kd texture, Ns in alpha
    mats[8] = mat_row(GGXREFL, (0.8, 0.7, 0.5), 0.2)
    mats = np.concatenate([mats, np.asarray([mat_row(GLASS, (0.04, 0.04, 0.04), ior=1.5),
                                             mat_row(BLINNPHONG, (0.2, 0.3, 0.6), 200.0)], np.float32)])
    sv, sf = _icosphere80()
    extra_p, extra_n = [], []
    for (c, r, m) in (((40.0, 30.0, 30.0), 30.0, 9), ((-50.0, 25.0, 60.0), 25.0, 10)):
        for f in sf:
            extra_p.append((sv[f] * r + np.asarray(c)).reshape(-1))
            extra_n.append(sv[f].reshape(-1))
        mat = np.concatenate([mat, np.full(len(sf), m, np.int32)])
    p = np.concatenate([p, np.asarray(extra_p, np.float32)])
    n = np.concatenate([n, np.asarray(extra_n, np.float32)])
    uv = np.concatenate([uv, np.zeros((len(extra_p), 6), np.float32)])
    return dict(base, p=p.astype(np.float32), n=n.astype(np.float32), uv=uv.astype(np.float32),
                mat=mat.astype(np.int32), mats=mats.astype(np.float32),
                mats_ext={1: [0.5, 0.5, 0.5, -1], 10: [0.9, 0.9, 0.9, 1]}, name="material_zoo_ext")
