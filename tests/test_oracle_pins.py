"""Pins of the oracle (the CPU restatement of the reference's ptnee path) that do not rely on the
oracle itself: published known-answer vectors, closed forms, invariants, brute force, and an
independent NumPy restatement of the SAH split rule.  CPU only."""
import math

import numpy as np
import pytest


# ------------------------------------------------------------------ Philox4x32-10
# Random123 known-answer vectors (kat_vectors, philox4x32 10 rounds).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(po, ctr, key, out):
    assert po.philox(ctr, key).tolist() == out


# ------------------------------------------------------------------ texture decode (texture.cpp:18-20)
def test_srgb8_decode_closed_form(po):
    vals = np.array([po.srgb8_decode(c) for c in range(256)], np.float64)
    assert vals[0] == 0.0 and vals[255] == 1.0
    assert np.all(np.diff(vals) > 0)
    exact = (np.arange(256) / 255.0) ** 2.2
    assert np.allclose(vals[1:], exact[1:], rtol=2e-7)  # exponent is the float literal 2.2f


# ------------------------------------------------------------------ triangle test (triangle.cpp:27-33)
def test_triangle_closed_form_and_bias(po):
    p = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    t, b1, b2 = po.tri_test(p, [0.25, 0.25, 1.0], [0, 0, -1])
    # s1.e1 = 1, so q = 1 + 1e-6: every output is scaled by 1 / (1 + 1e-6) (the reference's bias)
    k = 1.0 / (1.0 + 1e-6)
    assert abs(t - k) < 2e-7 and t < 1.0
    assert abs(b1 - 0.25 * k) < 1e-7 and abs(b2 - 0.25 * k) < 1e-7
    # double-sided: the reversed winding hits too, with q = -1 + 1e-6
    t2, c1, c2 = po.tri_test([0, 0, 0, 0, 1, 0, 1, 0, 0], [0.25, 0.25, 1.0], [0, 0, -1])
    k2 = -1.0 / (-1.0 + 1e-6)
    assert abs(t2 - k2) < 2e-7 and t2 > 1.0
    # edge: b1 = 0 exactly on the edge p0-p2
    _, e1, _ = po.tri_test(p, [0.0, 0.5, 1.0], [0, 0, -1])
    assert e1 == 0.0


def test_triangle_bias_grows_for_small_triangles(po):
    """With |e1 x e2| ~ 1e-6 the +1e-6 term halves t: the quirk is kept, not 'fixed'."""
    s = 1e-3
    t, _, _ = po.tri_test([0, 0, 0, s, 0, 0, 0, s, 0], [s / 4, s / 4, 1.0], [0, 0, -1])
    assert abs(t - 0.5) < 1e-4


# ------------------------------------------------------------------ slab test (bvhnode.cpp:35-48)
def test_slab_cases(po):
    mn, mx = [0, 0, 0], [1, 1, 1]
    assert po.slab(mn, mx, [-1, 0.5, 0.5], [1, 0, 0])          # zero components get +1e-9
    assert not po.slab(mn, mx, [-1, 2.0, 0.5], [1, 0, 0])      # parallel, outside the y slab
    assert po.slab(mn, mx, [0.5, 0.5, 0.5], [0, 0, 1])         # origin inside
    assert not po.slab(mn, mx, [0.5, 0.5, 2.0], [0, 0, 1])     # box behind: t_out < 0
    assert po.slab(mn, mx, [0.5, 0.5, 2.0], [0, 0, -1])
    # d = -1e-9 makes d + 1e-9 == 0: +-inf quotients, and 0/0 = NaN when the origin lies on the
    # bound; std::max / std::min keep the first operand on NaN, so the slab is then ignored
    assert po.slab(mn, mx, [0.5, 0.5, 0.5], [-1e-9, 0, 1])
    assert po.slab(mn, mx, [0.0, 0.5, 0.5], [-1e-9, 0, 1])
    assert not po.slab(mn, mx, [2.0, 0.5, 0.5], [-1e-9, 0, 1])


# ------------------------------------------------------------------ BVH build (bvh.cpp:5-97)
def _random_tris(n, seed, scale=100.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(0, scale, (n, 1, 3))
    return (c + rng.normal(0, scale / 30, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def _scene(p):
    n = len(p)
    return dict(p=p, n=None, uv=None, mat=np.zeros(n, np.int32),
                mats=np.array([[0, 0.5, 0.5, 0.5, 0.5, 1, -1, 0]], np.float32))


def _subtree_prims(ex, i):
    out = []
    stack = [i]
    while stack:
        v = stack.pop()
        if ex["left"][v] < 0:
            f, c = ex["leaf_first"][v], ex["leaf_count"][v]
            out.extend(ex["prim_order"][f:f + c].tolist())
        else:
            stack += [ex["right"][v], ex["left"][v]]
    return out


def _check_invariants(ex, p):
    n = len(p)
    order = ex["prim_order"]
    assert sorted(order.tolist()) == list(range(n))  # every triangle in exactly one leaf
    P = p.reshape(n, 3, 3)
    for v in range(len(ex["left"])):
        if ex["left"][v] < 0:
            f, c = ex["leaf_first"][v], ex["leaf_count"][v]
            assert 1 <= c <= 5 or n == 0
            pts = P[order[f:f + c]].reshape(-1, 3)
            assert np.array_equal(ex["box"][v, :3], pts.min(0)) and np.array_equal(ex["box"][v, 3:], pts.max(0))
        else:
            l, r = ex["left"][v], ex["right"][v]
            assert np.array_equal(ex["box"][v, :3], np.minimum(ex["box"][l, :3], ex["box"][r, :3]))
            assert np.array_equal(ex["box"][v, 3:], np.maximum(ex["box"][l, 3:], ex["box"][r, 3:]))


def _sah_split_numpy(prims, P, right_offset=0):
    """Independent restatement of BVH::_divide for inputs without pMax ties: returns the set of
    primitives that go left.  float32 arithmetic, same operation order as bvh.cpp:18-51; the
    right count is (n - i) (bvh.cpp:44), or (n - i - right_offset) for the self-check below."""
    prims = np.asarray(prims)
    pmin = P[prims].min(1)
    pmax = P[prims].max(1)
    n = len(prims)
    best = (np.float32(1e38), 0, 0)
    f32 = np.float32

    def sa(lo, hi):
        d = (hi - lo).astype(np.float32)
        return (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]) * f32(2)

    for axis in range(3):
        o = np.argsort(pmax[:, axis], kind="stable")
        lo_s = np.minimum.accumulate(pmin[o][::-1], 0)[::-1]
        hi_s = np.maximum.accumulate(pmax[o][::-1], 0)[::-1]
        lo_p = np.minimum.accumulate(pmin[o], 0)
        hi_p = np.maximum.accumulate(pmax[o], 0)
        for i in range(n - 1):
            cost = sa(lo_p[i], hi_p[i]) * f32(i + 1) + sa(lo_s[i + 1], hi_s[i + 1]) * f32(n - i - right_offset)
            if cost < best[0]:
                best = (cost, i + 1, axis)
    _, mid, axis = best
    o = np.argsort(pmax[:, axis], kind="stable")
    return set(prims[o[:mid]].tolist())


@pytest.mark.parametrize("n,seed", [(7, 1), (40, 2), (300, 3)])
def test_sah_split_rule_independent(po, n, seed):
    """Random triangles have distinct pMax values, so std::sort's order is unique and every split
    of the oracle's tree must equal the NumPy restatement (incl. the (n - i) right count)."""
    p = _random_tris(n, seed)
    assert all(len(np.unique(p.reshape(n, 3, 3).max(1)[:, a])) == n for a in range(3))
    ex = po.Scene(_scene(p)).export()
    _check_invariants(ex, p)
    P = p.reshape(n, 3, 3)
    for v in range(len(ex["left"])):
        if ex["left"][v] < 0:
            continue
        prims = _subtree_prims(ex, v)
        assert len(prims) > 5
        assert set(_subtree_prims(ex, ex["left"][v])) == _sah_split_numpy(prims, P)


def test_sah_pin_detects_off_by_one(po):
    """Self-check of the pin's power: the plausible slip (n - i - 1) for the right count
    (bvh.cpp:44) changes some splits of this tree, so the test above would catch it."""
    p = _random_tris(300, 3)
    n = len(p)
    P = p.reshape(n, 3, 3)
    ex = po.Scene(_scene(p)).export()
    differ = 0
    for v in range(len(ex["left"])):
        if ex["left"][v] < 0:
            continue
        prims = _subtree_prims(ex, v)
        if set(_subtree_prims(ex, ex["left"][v])) != _sah_split_numpy(prims, P, right_offset=1):
            differ += 1
    assert differ > 0


def test_box_room_tree_invariants(po, scenes):
    sc = scenes.box_room()
    ex = po.Scene(sc).export()
    _check_invariants(ex, sc["p"])


# ------------------------------------------------------------------ traversal (bvh.cpp:99-137)
def test_traversal_equals_brute_force(po):
    """The reference's brute-force loop (renderer.cpp:12-25) finds the same t on generic rays."""
    p = _random_tris(2000, 5)
    o = po.Scene(_scene(p))
    rng = np.random.Generator(np.random.PCG64(6))
    org = rng.uniform(-20, 120, (20000, 3)).astype(np.float32)
    d = rng.normal(size=(20000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    a = o.intersect(org, d)
    b = o.intersect(org, d, brute=True)
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[0], b[0])


def test_traversal_tie_order(po):
    """Duplicated triangles tie exactly; the winner is the one in the later DFS leaf, and within a
    leaf the first (bvh.cpp:119-126, bvhnode.cpp:57)."""
    base = _random_tris(30, 8)
    p = np.concatenate([base, base, base])
    o = po.Scene(_scene(p))
    ex = o.export()
    pos = np.empty(len(p), np.int64)
    leaf_of = np.empty(len(p), np.int64)
    for v in range(len(ex["left"])):
        if ex["left"][v] < 0:
            f, c = ex["leaf_first"][v], ex["leaf_count"][v]
            for k in range(c):
                pos[ex["prim_order"][f + k]] = f + k
                leaf_of[ex["prim_order"][f + k]] = f
    P = p.reshape(-1, 3, 3).astype(np.float64)
    cen = P.mean(1)
    org = (cen + np.array([0.0, 0.0, 500.0])).astype(np.float32)
    d = np.tile(np.array([0, 0, -1], np.float32), (len(p), 1))
    tri, t, _, _ = o.intersect(org, d)
    bt, bb, _, _ = o.intersect(org, d, brute=True)
    assert np.array_equal(t.view(np.uint32), bb.view(np.uint32))
    for i in range(len(p)):
        if tri[i] < 0:
            continue
        # candidates with the same t: the three copies of the winning triangle at least
        same = [j for j in range(len(p)) if j % 30 == tri[i] % 30]
        best = max(same, key=lambda j: (leaf_of[j], -pos[j]))
        assert tri[i] == best


# ------------------------------------------------------------------ camera (camera.cpp:24-32,88-105)
def test_camera_closed_form(po):
    W, H = 320.0, 240.0
    cam = np.array([0, 100, 350, 0, 0, 0, 40, 4 / 3], np.float32)
    d = po.camera_rays(cam, [160.0, 0.0, 320.0], [120.0, 0.0, 240.0], W, H)
    assert np.allclose(d[0], [0, 0, -1], atol=1e-7)
    fh = 2 * math.tan(math.radians(40) / 2)
    fw = fh * 4 / 3
    c = np.array([-fw / 2, fh / 2, -1.0]); c /= np.linalg.norm(c)
    assert np.allclose(d[1], c, atol=2e-7)                 # top-left: y goes down the image
    c = np.array([fw / 2, -fh / 2, -1.0]); c /= np.linalg.norm(c)
    assert np.allclose(d[2], c, atol=2e-7)
    # yaw 90: gaze = (-1, 0, 0); pitch 30: gaze.y = sin 30
    g = po.camera_rays(np.array([0, 0, 0, 90, 0, 0, 40, 1], np.float32), [0.5], [0.5], 1, 1)[0]
    assert np.allclose(g, [-1, 0, 0], atol=1e-5)
    g = po.camera_rays(np.array([0, 0, 0, 0, 30, 0, 40, 1], np.float32), [0.5], [0.5], 1, 1)[0]
    assert abs(g[1] - 0.5) < 1e-5 and g[2] < 0


# ------------------------------------------------------------------ materials
def _mat_scene(po, rows, ext=None):
    p = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    return po.Scene(dict(p=p, n=None, uv=None, mat=np.zeros(1, np.int32), mats=np.array(rows, np.float32),
                         mats_ext=ext), build_bvh=False)


def test_lambert_f_pdf_and_sampling(po):
    """MatLambert (matlambert.cpp): f = Kd/pi on the same side, 0 across; pdf = (|cos| + 1e-4)/pi;
    sampling has cos^2 uniform on [0.01, 1] (r^2 = 0.99 U)."""
    s = _mat_scene(po, [[0, 0.5, 0.25, 0.125, 0.5, 1, -1, 0]])
    n = np.tile([0, 0, 1.0], (4, 1))
    wo = np.tile([0, 0.6, 0.8], (4, 1))
    wi = np.array([[0, 0, 1.0], [0.6, 0, 0.8], [0, 0, -1.0], [0.8, 0, -0.6]])
    _, f, pdf = s.material_eval(0, wo, n, wi=wi)
    pi = 3.14159
    assert np.allclose(f[:2], np.array([0.5, 0.25, 0.125]) / pi, rtol=1e-6) and not f[2:].any()
    assert np.allclose(pdf, (np.abs(wi[:, 2]) + 1e-4) / pi, rtol=1e-6)
    rng = np.random.Generator(np.random.PCG64(1))
    k = 200000
    u = rng.random((k, 3)).astype(np.float32)
    w, _, _ = s.material_eval(0, np.tile([0, 0.6, 0.8], (k, 1)), np.tile([0, 0, 1.0], (k, 1)), u3=u)
    c2 = np.sort(w[:, 2].astype(np.float64) ** 2)
    assert abs(c2.min() - 0.01) < 1e-3 and c2.max() <= 1.0 + 1e-6
    ks = np.abs(c2 - (0.01 + 0.99 * (np.arange(k) + 0.5) / k)).max()
    assert ks < 0.01


def test_ggx_ndf_normalised_and_sampling(po):
    """MatGGXRefl: D integrates to 1 against cos(theta_h); pdf = D (n.h)/4 + 1e-8 away from the
    horizon, 1e18 below it; sampled half vectors follow tan^2(theta_h)/alpha^2 = U/(1-U)."""
    a = 0.3
    th = np.linspace(0, np.pi / 2, 200001)
    c = np.cos(th)
    D = a * a / 3.14159 / (c * c * (a * a - 1) + 1) ** 2
    integral = np.trapezoid(D * c * np.sin(th), th) * 2 * np.pi
    assert abs(integral - 3.14159 / np.pi) < 1e-4   # the reference's pi is 3.14159
    s = _mat_scene(po, [[2, 0.9, 0.9, 0.9, a, 1, -1, 0]])
    wo = np.tile([0.0, 0.6, 0.8], (3, 1))
    n = np.tile([0.0, 0.0, 1.0], (3, 1))
    wi = np.array([[0.0, -0.6, 0.8], [0.6, 0.0, 0.8], [0, 0.6, -0.8]])
    _, f, pdf = s.material_eval(0, wo, n, wi=wi)
    h = (wo[0] + wi[0]) / np.linalg.norm(wo[0] + wi[0])
    ch = h[2]
    d0 = a * a / 3.14159 / (ch * ch * (a * a - 1) + 1) ** 2
    assert abs(pdf[0] - (d0 * ch / 4 + 1e-8)) / pdf[0] < 1e-5
    assert pdf[2] == np.float32(1e18) and not f[2].any()
    rng = np.random.Generator(np.random.PCG64(2))
    k = 100000
    u = rng.random((k, 3)).astype(np.float32)
    wos = np.tile([0.0, 0.0, 1.0], (k, 1))
    w, _, _ = s.material_eval(0, wos, np.tile([0, 0, 1.0], (k, 1)), u3=u)
    hv = w.astype(np.float64) + wos
    hv /= np.linalg.norm(hv, axis=1, keepdims=True)
    t2 = np.sort((1 - hv[:, 2] ** 2) / hv[:, 2] ** 2 / a ** 2)
    x = t2 / (1 + t2)  # should be ~U (theta scaled by 0.9999)
    assert np.abs(x - (np.arange(k) + 0.5) / k).max() < 0.01


def test_blinnphong_f_pdf_and_sampling(po):
    """MatBlinnPhong (matblinnphong.cpp): f = Kd/pi + (Ns+2)/(8 pi) Ks (n.h)^Ns above the horizon;
    pdf = (|cos|+1e-6)/(2 pi) + (Ns+2)/(16 pi) (n.h)^(Ns+1), 1e-18 below it; the diffuse lobe (u0 <
    0.5) has cos^2 uniform on [0.01, 1]; the glossy lobe has cos_h = r1^(1/(Ns+2)) with the
    reference's sin_h = sqrt(1 - cos_h), wi = 2 (h.wo) h - wo left unnormalised."""
    ns = 20.0
    kd, ks = np.array([0.5, 0.25, 0.125]), np.array([0.3, 0.6, 0.9])
    s = _mat_scene(po, [[4, *kd, ns, 1, -1, 0]], ext={0: [*ks, -1]})
    n = np.tile([0, 0, 1.0], (3, 1))
    wo = np.tile([0, 0.6, 0.8], (3, 1))
    wi = np.array([[0.6, 0, 0.8], [0, -0.6, 0.8], [0.8, 0, -0.6]])
    _, f, pdf = s.material_eval(0, wo, n, wi=wi)
    pi = 3.14159
    for i in range(2):
        h = (wo[i] + wi[i]) / np.linalg.norm(wo[i] + wi[i])
        spec = (ns + 2) / pi / 8 * h[2] ** ns
        assert np.allclose(f[i], kd / pi + ks * spec, rtol=1e-5)
        want = 0.5 * (abs(wi[i, 2]) + 1e-6) / pi + 0.5 * (ns + 2) / pi / 8 * h[2] ** (ns + 1)
        assert abs(pdf[i] - want) / want < 1e-5
    assert not f[2].any() and pdf[2] == np.float32(1e-18)
    rng = np.random.Generator(np.random.PCG64(3))
    k = 200000
    u = rng.random((k, 3)).astype(np.float32)
    w, _, _ = s.material_eval(0, np.tile([0, 0, 1.0], (k, 1)), np.tile([0, 0, 1.0], (k, 1)), u3=u)
    w = w.astype(np.float64)
    diff = u[:, 0] < 0.5
    c2 = np.sort(w[diff, 2] ** 2)
    m = len(c2)
    assert np.abs(c2 - (0.01 + 0.99 * (np.arange(m) + 0.5) / m)).max() < 0.01
    assert np.allclose(np.linalg.norm(w[diff], axis=1), 1, atol=1e-5)
    # glossy lobe with wo = n: wi.z = 2 c^2 - 1 and |wi_xy| = 2 c sqrt(1 - c), c = cos_h
    g = w[~diff]
    ct = u[~diff, 1].astype(np.float64) ** (1 / (ns + 2))
    assert np.allclose(g[:, 2], 2 * ct * ct - 1, atol=1e-6)
    assert np.allclose(np.hypot(g[:, 0], g[:, 1]), 2 * ct * np.sqrt(1 - ct), atol=1e-6)
    x = np.sort(((g[:, 2] + 1) / 2) ** ((ns + 2) / 2))   # cos_h^(Ns+2) read back from wi: uniform
    m = len(x)
    assert np.abs(x - (np.arange(m) + 0.5) / m).max() < 0.01


def test_glass_snell_fresnel_and_half_pdf(po):
    """MatGlass (matglass.cpp): refraction with sin_i = sin_o / ior entering (ior outside-in
    swapped when leaving, clamped at 1), mirror reflection, Schlick F = F0 + (1-F0)(1-|cos_o|)^5
    for reflection, (1-F) eta^2 for transmission, pdf 1/2, f = 0 within 1e-3 of grazing."""
    f0, ior = 0.04, 1.5
    s = _mat_scene(po, [[5, f0, f0, f0, 0.5, ior, -1, 0]])
    so = 0.6
    wo = np.array([[so, 0, 0.8], [so, 0, 0.8], [so, 0, -0.8], [0.9, 0, -np.sqrt(1 - 0.81)]])
    n = np.tile([0, 0, 1.0], (4, 1))
    u = np.array([[0.2, 0, 0], [0.7, 0, 0], [0.2, 0, 0], [0.2, 0, 0]], np.float32)
    w, _, pdf = s.material_eval(0, wo, n, u3=u)
    si = so / ior
    assert np.allclose(w[0], [-si, 0, -np.sqrt(1 - si * si)], atol=1e-6)     # refracted inwards
    assert np.allclose(w[1], [-so, 0, 0.8], atol=1e-6)                     # mirror
    sx = min(so * ior, 1.0)
    assert np.allclose(w[2], [-sx, 0, np.sqrt(1 - sx * sx)], atol=1e-6)     # leaving the glass
    assert np.allclose(w[3], [-1, 0, 0], atol=1e-6)                        # total internal: grazing
    assert np.all(pdf == 0.5)
    wi = np.array([[-si, 0, -np.sqrt(1 - si * si)], [-so, 0, 0.8], [-0.999, 0, np.sqrt(1 - 0.999 ** 2)]])
    _, f, _ = s.material_eval(0, np.tile(wo[0], (3, 1)), n[:3], wi=wi)
    F = f0 + (1 - f0) * (1 - 0.8) ** 5
    assert np.allclose(f[0], (1 - F) / ior ** 2, rtol=1e-5)
    assert np.allclose(f[1], F, rtol=1e-5)
    assert np.allclose(f[2], F, rtol=1e-5)   # |cos_i| = 0.045 > 1e-3: still reflects
    _, f, _ = s.material_eval(0, wo[2:3], n[:1], wi=np.array([[-0.6, 0, 0.8]]))   # leaving: eta = ior
    assert np.allclose(f[0], (1 - F) * ior ** 2, rtol=1e-5)
    _, f, _ = s.material_eval(0, wo[:1], n[:1], wi=np.array([[-1, 0, 5e-4]]))
    assert not f.any()


def test_light_sampler_area_weights_and_2x_pdf(po, scenes):
    """LightSampler: CDF over emitters by area() = |e1 x e2| (twice the area), p() = 1/sum."""
    sc = scenes.box_room()
    o = po.Scene(sc, build_bvh=False)
    assert abs(o.light_p() - 1.0 / 5000.0) < 1e-9   # 50 x 50 light: two triangles of |e1 x e2| = 2500
    lights = np.nonzero(sc["mats"][sc["mat"]][:, 0] == 1)[0]
    picks = [o.light_pick(u) for u in np.linspace(0.001, 0.999, 999)]
    assert set(picks) == set(lights.tolist())
    assert abs(np.mean(np.array(picks) == lights[0]) - 0.5) < 0.01


def test_texture_bilinear_black_border(po):
    """Texture::pixelUV / pixelBi (texture.cpp:54-72): bilinear between floor and floor+1, texels
    outside the image read black (no wrap of the +1 neighbour), u and 1-v wrapped into [0, 1)."""
    img = np.zeros((2, 2, 4), np.uint8)
    img[0, 0, :3] = 255   # (x0, y0)
    img[0, 1, 0] = 255    # (x1, y0) red
    sc = dict(p=np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), n=None, uv=None, mat=np.zeros(1, np.int32),
              mats=np.array([[0, 1, 1, 1, 0.5, 1, 0, 0]], np.float32), textures=[img])
    o = po.Scene(sc, build_bvh=False)
    # v is flipped: v = 1 -> row 0.  (u, v) = (0, 1) -> x = 0, y = 0: exactly texel (0, 0)
    out = o.texture_lookup(0, [0.0, 0.25, 0.5, 0.75, 1.25], [1.0, 1.0, 1.0, 1.0, 1.0])
    assert np.allclose(out[0], [1, 1, 1])
    assert np.allclose(out[1], [1, 0.5, 0.5])          # x = 0.5: half white + half red
    assert np.allclose(out[2], [1, 0, 0])              # x = 1: texel (1, 0)
    assert np.allclose(out[3], [0.5, 0, 0])            # x = 1.5: red fades into the black border
    assert np.allclose(out[4], out[1])                 # u wraps


# ------------------------------------------------------------------ estimator closed forms
def _plane_scene(albedo, env_value, env_w=1024):
    R = 1e4
    p = np.array([[-R, 0, R, R, 0, R, R, 0, -R], [R, 0, -R, -R, 0, -R, -R, 0, R]], np.float32)
    env = np.full((8, env_w, 3), env_value, np.float32)
    return dict(p=p, n=None, uv=None, mat=np.zeros(2, np.int32),
                mats=np.array([[0, albedo, albedo, albedo, 0.5, 1, -1, 0]], np.float32),
                envmap=dict(rgb=env, gain=1.0))


def test_lambert_plane_under_constant_sky(po):
    """A Lambert plane of albedo rho under a constant environment E, seen from above: the camera
    ray hits the plane, the BSDF-sampled ray escapes to the sky.  With no emitters there is no
    NEE; RR divides by prr.  The expectation is E rho E[c/(c+1e-4)] with c^2 ~ U[0.01, 1]
    (matlambert.cpp:7,36), times (1 - 0.5/W) for the black border at the u = 1 seam of a W-wide
    environment texture (texture.cpp:54-67)."""
    rho, E, Wenv = 0.6, 2.0, 1024
    o = po.Scene(_plane_scene(rho, E, Wenv))
    cam = np.array([0, 10, 0, 0, -90, 0, 10, 1], np.float32)
    spp = 20000
    s, s2 = o.render(cam, 4, 4, spp, seed=3)
    cg = np.linspace(0.1, 1, 2000001)
    ratio = np.trapezoid(2 * cg / 0.99 * cg / (cg + 1e-4), cg)
    expect = E * rho * ratio * (1 - 0.5 / Wenv)
    N = 16 * spp
    mean = s.sum() / (3 * N)
    var = s2.sum() / (3 * N) - mean ** 2
    z = (mean - expect) / math.sqrt(var / N)
    assert abs(z) < 5, (mean, expect)


# ------------------------------------------------------------------ sky sampling (optional, not in the reference)
def _sky_plane_scene(mat_row, sky):
    R = 1e4
    p = np.array([[-R, 0, R, R, 0, R, R, 0, -R], [R, 0, -R, -R, 0, -R, -R, 0, R]], np.float32)
    return dict(p=p, n=None, uv=None, mat=np.zeros(2, np.int32), mats=np.array([mat_row], np.float32),
                envmap=dict(rgb=sky, gain=1.0))


def _sphere_grid(nt=1500, nph=3000):
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi - np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).reshape(-1)
    return d, dw


def test_env_sampler_density_normalised_and_matches_its_samples(po, scenes):
    """The sky sampler's density integrates to 1 over the sphere (quadrature), equals the density
    reported with each sample, and its cell frequencies follow the cell probabilities
    (max corner texel x sin theta, normalised) to 4 standard errors."""
    sky = scenes._sky(64, 32)
    o = po.Scene(_sky_plane_scene([0, .5, .5, .5, .5, 1, -1, 0], sky), build_bvh=False)
    d, dw = _sphere_grid()
    assert abs((o.env_density(d) * dw).sum() - 1.0) < 2e-3
    rng = np.random.Generator(np.random.PCG64(1))
    n = 400000
    ds, ps, ok = o.env_sample(rng.random((n, 4)).astype(np.float32))
    assert ok.mean() > 0.9999
    assert np.allclose(o.env_density(ds[ok].astype(np.float64)), ps[ok], rtol=1e-4)
    # cell probabilities from the texels, independently of the sampler's tables
    H, W = sky.shape[:2]
    m = np.abs(sky).max(-1)
    mp = np.zeros((H + 1, W + 1)); mp[:H, :W] = m
    c = np.maximum(np.maximum(mp[:-1, :-1], mp[:-1, 1:]), np.maximum(mp[1:, :-1], mp[1:, 1:]))
    wgt = c * np.sin((np.arange(H) + 0.5) / H * np.float32(3.14159))[:, None]
    prob = wgt / wgt.sum()
    th = np.arctan2(np.hypot(ds[ok, 0], ds[ok, 2]), ds[ok, 1])
    ph = np.arctan2(ds[ok, 2], ds[ok, 0])
    i = np.clip(np.floor(th / np.float32(3.14159) * H).astype(int), 0, H - 1)
    j = np.clip(np.floor((ph / (2 * np.float32(3.14159)) + 0.5) * W).astype(int), 0, W - 1)
    cnt = np.bincount(i * W + j, minlength=H * W).reshape(H, W)
    exp = prob * ok.sum()
    sel = exp > 20
    z = (cnt[sel] - exp[sel]) / np.sqrt(exp[sel])
    assert np.abs(z).max() < 4.5 and sel.sum() > 500


def test_bsdf_sampler_true_densities(po):
    """The true densities of the Lambert and GGX-reflect samplers used by sky sampling: Lambert
    q = cos/(0.99 pi) above cos = 0.1 integrates to 1; GGX-reflect q integrates to 1 for wo = n
    and agrees with a histogram of the sampler's own directions (binned in cos theta_i) for an
    oblique wo; below the horizon its missing mass is where the BSDF is zero."""
    s = _mat_scene(po, [[0, .5, .5, .5, .5, 1, -1, 0], [2, .8, .8, .8, 0.3, 1, -1, 0]])
    d, dw = _sphere_grid()
    dz = d[:, [0, 2, 1]]   # grid around +z
    nz = np.tile([0, 0, 1.0], (len(d), 1))
    q0 = s.sample_density(0, np.tile([0, .6, .8], (len(d), 1)), nz, dz)
    assert abs((q0 * dw).sum() - 1) < 1e-3
    c = dz[:, 2]
    assert np.allclose(q0, np.where(c > 0.1, c / (np.float32(0.99) * np.pi), 0), rtol=1e-6, atol=1e-12)
    q1 = s.sample_density(1, np.tile([0, 0, 1.0], (len(d), 1)), nz, dz)
    assert abs((q1 * dw).sum() - 1) < 2e-3
    wo = np.array([0, 0.6, 0.8])
    q1o = s.sample_density(1, np.tile(wo, (len(d), 1)), nz, dz)
    rng = np.random.Generator(np.random.PCG64(4))
    k = 400000
    w, _, _ = s.material_eval(1, np.tile(wo, (k, 1)), np.tile([0, 0, 1.0], (k, 1)),
                              u3=rng.random((k, 3)).astype(np.float32))
    edges = np.linspace(0.05, 1, 20)
    hist = np.histogram(w[:, 2].astype(np.float64), bins=edges)[0] / k
    want = np.array([(q1o * dw)[(c >= a) & (c < b)].sum() for a, b in zip(edges[:-1], edges[1:])])
    assert np.abs(hist - want).max() < 4e-3


def test_env_sampling_keeps_the_expected_image(po, scenes):
    """Sky sampling is a variance reduction of the reference's estimator: the Lambert plane under a
    constant sky keeps its closed form, and under the sun-lobe sky the images with and without it
    agree per pixel (z-test) for Lambert and GGX-reflect planes while the Lambert variance drops."""
    rho, E, Wenv = 0.6, 2.0, 1024
    o = po.Scene(_plane_scene(rho, E, Wenv))
    spp = 20000
    s, s2 = o.render(np.array([0, 10, 0, 0, -90, 0, 10, 1], np.float32), 4, 4, spp, seed=3, env_is=True)
    cg = np.linspace(0.1, 1, 2000001)
    expect = E * rho * np.trapezoid(2 * cg / 0.99 * cg / (cg + 1e-4), cg) * (1 - 0.5 / Wenv)
    N = 16 * spp
    mean = s.sum() / (3 * N)
    assert abs(mean - expect) < 5 * math.sqrt((s2.sum() / (3 * N) - mean ** 2) / N)
    sky = scenes._sky(128, 64)
    cam = np.array([0, 10, 0, 30, -60, 0, 40, 1], np.float32)
    zs = []
    for row in ([0, .6, .6, .6, .5, 1, -1, 0], [2, .8, .7, .6, 0.3, 1, -1, 0]):
        res = []
        for env_is in (0, 1):
            spp = 3000
            s, s2 = po.Scene(_sky_plane_scene(row, sky)).render(cam, 12, 12, spp, seed=7 + env_is, env_is=env_is)
            m = s / spp
            res.append((m, np.maximum(s2 / spp - m * m, 0) / spp))
        (m0, v0), (m1, v1) = res
        z = (m1 - m0) / np.sqrt(v0 + v1)
        zs.append(z)
        assert (np.abs(z) > 4).sum() == 0
        if row[0] == 0:
            assert v0.mean() > 5 * v1.mean()
    z = np.concatenate(zs)
    assert abs(z.mean()) < 4 / math.sqrt(len(z))


def light_over_plane(rho=0.5, Le=10.0, a=2.0, h=10.0):
    """A Lambert plane (albedo rho, +-1e4) under an a x a emitter of radiance Le facing down at
    height h; returns the scene, a 1-pixel camera looking straight down at the centre, and the
    irradiance Ed at the centre (numerical quadrature of the square's form factor)."""
    R = 1e4
    p = np.array([[-R, 0, R, R, 0, R, R, 0, -R], [R, 0, -R, -R, 0, -R, -R, 0, R],
                  [-a / 2, h, -a / 2, a / 2, h, -a / 2, a / 2, h, a / 2],
                  [a / 2, h, a / 2, -a / 2, h, a / 2, -a / 2, h, -a / 2]], np.float32)
    # light faces down: (p1 - p0) x (p2 - p0) must point to -y
    e = np.cross(p[2, 3:6] - p[2, :3], p[2, 6:] - p[2, :3])
    if e[1] > 0:
        p[2] = np.concatenate([p[2, :3], p[2, 6:], p[2, 3:6]])
        p[3] = np.concatenate([p[3, :3], p[3, 6:], p[3, 3:6]])
    sc = dict(p=p, n=None, uv=None, mat=np.array([0, 0, 1, 1], np.int32),
              mats=np.array([[0, rho, rho, rho, 0.5, 1, -1, 0], [1, Le, Le, Le, 0.5, 1, -1, 0]], np.float32))
    cam = np.array([0, h / 2, 0, 0, -90, 0, 0.01, 1], np.float32)
    x = np.linspace(-a / 2, a / 2, 1201)
    X, Y = np.meshgrid(x, x)
    Ed = Le * np.trapezoid(np.trapezoid(h * h / (X ** 2 + Y ** 2 + h * h) ** 2, x), x)
    return sc, cam, Ed


def test_nee_twice_weight_with_fixed_mis(po):
    """mis = 0 fixes the MIS weight to 0.5 (rendererptls.cpp:65-68).  With LightSampler::p() being
    half the true area pdf (triangle.cpp:57-60), the NEE term estimates twice the direct light,
    the BSDF-sampled emitter hit the direct light itself: the total is 1.5 x E_direct."""
    rho = 0.5
    sc, cam, Ed = light_over_plane(rho)
    o = po.Scene(sc)
    spp = 100000
    s, s2 = o.render(cam, 1, 1, spp, seed=4, mis=0, maxdepth=2)
    m = s[0] / spp
    # BSDF term: Lambert samples cos >= 0.1 with density cos/(0.99 pi) but reports (cos + 1e-4)/pi,
    # so its expectation carries a factor 1/0.99 (the light subtends cos ~ 1: 1e-4 is negligible)
    expect = (0.5 * 2.0 + 0.5 / 0.99) * rho / 3.14159 * Ed
    se = math.sqrt(s2[0] / spp - m * m) / math.sqrt(spp)
    assert abs(m - expect) < 5 * se + 1e-3 * expect, (m, expect, se)


def test_pt_direct_light_closed_form(po):
    """The pt integrator (rendererpt.cpp:13-57) has no light sampling: the plane's radiance is the
    BSDF-sampled emitter hit alone, rho/(0.99 pi) E_direct (the Lambert pdf's 1/0.99 as above), with
    Russian roulette from the first hit unbiased."""
    rho = 0.5
    sc, cam, Ed = light_over_plane(rho, a=10.0, h=5.0)   # subtends cos >= 0.57: inside the sampled cone
    spp = 200000
    s, s2 = po.Scene(sc).render(cam, 1, 1, spp, seed=5, integrator=1)
    m = s[0] / spp
    expect = rho / (0.99 * 3.14159) * Ed
    se = math.sqrt(s2[0] / spp - m * m) / math.sqrt(spp)
    assert abs(m - expect) < 5 * se + 1e-3 * expect, (m, expect, se)
    assert 5 * se < 0.1 * expect   # the test has the power to see a dropped 1/prr or a doubled term


# ------------------------------------------------------------------ adjoint (FP64 gradient)
def test_adjoint_albedo_equals_crn_central_difference(po, scenes):
    """Albedo / F0 does not move any sampled direction and enters each sample polynomially, so with
    common random numbers the central difference of the FP64 render equals the gradient up to
    O(h^2): an independent check of every albedo term (emitter hit, NEE, continuation)."""
    sc, theta_star, theta0 = scenes.adjoint_problem(width=12, height=12)
    W, H, spp = 12, 12, 6
    M = len(theta0)
    cam = scenes.camera_array(sc["camera"])
    rng = np.random.Generator(np.random.PCG64(8))
    dldi = rng.normal(size=(H, W, 3))
    o = po.Scene(sc)
    o.set_params(theta0)
    g = o.adjoint(cam, W, H, spp, dldi, M, seed=21)

    def loss(theta):
        o.set_params(theta)
        s, _ = o.render(cam, W, H, spp, seed=21, fp64=True)
        return (dldi.reshape(-1) * s / spp).sum()
    h = 1e-4
    for mi, ch in [(0, 0), (9, 1), (23, 2), (44, 0), (58, 1), (63, 2)]:
        tp, tm = theta0.copy(), theta0.copy()
        tp[mi, ch] += h
        tm[mi, ch] -= h
        fd = (loss(tp) - loss(tm)) / (float(tp[mi, ch]) - float(tm[mi, ch]))   # float32 step taken
        assert abs(g[mi, ch] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (mi, ch, g[mi, ch], fd)


@pytest.mark.parametrize("integrator", [0, 1])
def test_adjoint_through_glass_and_blinnphong_equals_crn_difference(po, scenes, integrator):
    """The same check on paths that cross glass (refraction inside the sphere, back-face hits of a
    transmissive model) and Blinn-Phong (replayed, not differentiated), for ptnee and pt: the
    Lambert and GGX-reflect rows must equal the central difference of the FP64 render."""
    sc = scenes.material_zoo_ext(16, 12)
    W, H, spp = 16, 12, 6
    M = len(sc["mats"])
    theta0 = sc["mats"][:, 1:5].astype(np.float32).copy()
    cam = scenes.camera_array(sc["camera"])
    dldi = np.random.Generator(np.random.PCG64(12)).normal(size=(H, W, 3))
    o = po.Scene(sc)
    g = o.adjoint(cam, W, H, spp, dldi, M, seed=17, integrator=integrator)

    def loss(theta):
        o.set_params(theta)
        s, _ = o.render(cam, W, H, spp, seed=17, fp64=True, integrator=integrator)
        return (dldi.reshape(-1) * s / spp).sum()
    h = 1e-4
    for mi, ch in [(2, 0), (3, 1), (4, 2), (8, 0), (8, 2)]:
        tp, tm = theta0.copy(), theta0.copy()
        tp[mi, ch] += h
        tm[mi, ch] -= h
        fd = (loss(tp) - loss(tm)) / (float(tp[mi, ch]) - float(tm[mi, ch]))
        assert abs(g[mi, ch] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (mi, ch, g[mi, ch], fd)
    o.set_params(theta0)


def test_adjoint_roughness_expectation_on_ggx_plane(po):
    """d/d alpha of the mean image of a GGX-reflect plane under a constant sky, seen straight down
    (wo = n): the expectation is E(alpha) = Env (1 - 0.5/W) * integral over U of
    f cos / pdf at the half vector theta_h = 0.9999 atan(alpha sqrt(U/(1-U))) (matggxrefl.cpp:5-34),
    evaluated by quadrature from the material's f and pdf and differentiated by central difference.
    The adjoint's score-function estimate must agree within its standard error."""
    a, F0, Env, Wenv = 0.35, 0.7, 1.5, 1024
    R = 1e4
    p = np.array([[-R, 0, R, R, 0, R, R, 0, -R], [R, 0, -R, -R, 0, -R, -R, 0, R]], np.float32)
    sc = dict(p=p, n=None, uv=None, mat=np.zeros(2, np.int32),
              mats=np.array([[2, F0, F0, F0, a, 1, -1, 0]], np.float32),
              envmap=dict(rgb=np.full((8, Wenv, 3), Env, np.float32), gain=1.0))

    def expect(alpha):
        s = _mat_scene(po, [[2, F0, F0, F0, alpha, 1, -1, 0]])
        k = 400000
        u = (np.arange(k) + 0.5) / k
        th = 0.9999 * np.arctan(alpha * np.sqrt(u / (1 - u)))
        wi = np.stack([np.sin(2 * th), np.zeros(k), np.cos(2 * th)], 1)
        n = np.tile([0, 0, 1.0], (k, 1))
        _, f, pdf = s.material_eval(0, n, n, wi=wi)
        g = f[:, 0].astype(np.float64) * np.clip(wi[:, 2], 0, None) / pdf.astype(np.float64)
        g[wi[:, 2] <= 0] = 0
        return Env * (1 - 0.5 / Wenv) * g.mean()
    h = 2e-3
    dexp = (expect(a + h) - expect(a - h)) / (2 * h)
    o = po.Scene(sc)
    cam = np.array([0, 10, 0, 0, -90, 0, 0.01, 1], np.float32)
    W = H = 4
    dldi = np.full((H, W, 3), 1.0 / (W * H * 3))
    runs = [o.adjoint(cam, W, H, 4000, dldi, 1, seed=100 + s)[0, 3] for s in range(8)]
    m, se = np.mean(runs), np.std(runs, ddof=1) / math.sqrt(len(runs))
    assert abs(m - dexp) < 5 * se + 2e-3 * abs(dexp), (m, se, dexp)
    assert abs(dexp) > 20 * se   # the check has power: the derivative is well resolved


# ------------------------------------------------------------------ radiance-cache network
def test_nrc_encoding_known_values(po):
    """Triangle wave of renderernrc.cpp:50-53 at points with known values: tri(0) = 1,
    tri(1) = -1, tri(1/2) = 0, period 2, and C fmod's sign for negative x: tri(-1/2) = 2."""
    import nrcdata
    q = np.zeros((4, 16), np.float32)
    q[:, 0] = [0.0, 1.0, -0.5, 6.5]
    q[:, 3] = [0.0, 1.0, 0.5, 2.0]
    q[:, 4:8] = 0.25
    q[:, 8:11] = 0.7
    f = po.nrc_encode(q)
    assert list(f[:, 0]) == [1.0, -1.0, 2.0, 0.0]          # x / 1: 0, 1, -0.5, 6.5 (fmod 0.5)
    assert list(f[:, 1]) == [1.0, 0.0, 1.5, -0.5]          # x / 2: 0, 0.5, -0.25, 3.25 (fmod 1.25)
    assert np.all(f[:, 36:40] == 0.25) and np.all(f[:, 41:44] == np.float32(0.7))
    assert np.allclose(f[:, 40], np.exp(-q[:, 3]), rtol=1e-7)
    assert np.all(f[:, 44] == 1.0) and not f[:, 45:].any()
    assert np.abs(f[:, 0:36]).max() <= 5.0
    # every frequency of a positive coordinate lies in [-1, 1]
    g = po.nrc_encode(nrcdata.queries(1000, seed=3))
    assert np.abs(g[:, 0:36]).max() <= 1.0


def test_nrc_forward_identity_network(po):
    """Weights that route the albedo inputs straight to the outputs (W^T rows = output units):
    the exact network returns the albedo, the BF16 network its BF16 rounding."""
    import nrcdata
    q = nrcdata.queries(500, seed=1)
    w = nrcdata.identity_weights(po.NRC_NPARAM)
    y = po.nrc_forward(w, q)
    assert np.array_equal(y, q[:, 8:11].astype(np.float64))
    yb = po.nrc_forward(w, q, bf16=True)
    a = q[:, 8:11].astype(np.float32)
    ab = (a.view(np.uint32) + 0x7FFF + ((a.view(np.uint32) >> 16) & 1)) & 0xFFFF0000
    assert np.array_equal(yb, ab.view(np.float32).astype(np.float64))


def test_nrc_oracle_gradient_central_difference(po):
    """The FP64 network's loss gradient against central differences.  The relative L2 loss holds
    its denominator y^2 + 0.01 constant (stop-gradient), so the differenced function is
    mean sum (y(w) - t)^2 / (y(w0)^2 + 0.01) with y(w0) at the base weights."""
    import nrcdata
    rng = np.random.Generator(np.random.PCG64(4))
    w = np.zeros(po.NRC_NPARAM, np.float32)
    for l in range(6):
        rows = 64 if l < 5 else 3
        cols = 45 if l == 0 else 64
        n_rows = 64 if l < 5 else 16
        w[l * 4096: l * 4096 + n_rows * 64].reshape(n_rows, 64)[:rows, :cols] = \
            rng.uniform(-1, 1, (rows, cols)) * np.sqrt(6 / cols)
    q = nrcdata.queries(64, seed=2)
    t = nrcdata.smooth_target(q).astype(np.float64)
    loss, g = po.nrc_grad(w, q, t)
    y0 = po.nrc_forward(w, q)
    den = y0 ** 2 + 0.01
    assert abs(loss - ((y0 - t) ** 2 / den).sum(1).mean()) < 1e-12 * max(loss, 1.0)

    def surrogate(wv):
        y = po.nrc_forward(wv, q)
        return ((y - t) ** 2 / den).sum(1).mean()
    idx = [5, 64 * 3 + 44, 4096 + 70, 2 * 4096 + 64 * 10 + 9, 4 * 4096 + 64 * 63 + 1, 5 * 4096 + 2 * 64 + 17,
           5 * 4096 + 5]
    for i in idx:
        h = 1e-5 * max(1.0, abs(float(w[i])))   # small enough that no ReLU changes side
        wp, wm = w.copy(), w.copy()
        wp[i] = np.float32(float(w[i]) + h)
        wm[i] = np.float32(float(w[i]) - h)
        fd = (surrogate(wp) - surrogate(wm)) / (float(wp[i]) - float(wm[i]))
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, fd, g[i])
    assert g[5 * 4096 + 3 * 64 + 5] == 0.0 and g[45] == 0.0   # unused output row, zero input


def test_adjoint_texel_gradient_equals_crn_central_difference(po, scenes):
    """Texel values enter each sample through the bilinear albedo (texture.cpp:61-67), linearly in
    each texel, so the central difference of the FP64 render with common random numbers equals
    the texel gradient up to O(h^2) -- floor (Lambert) and back-wall (GGX-reflect F0) textures."""
    sc = scenes.textured_room(width=12, height=10)
    W, H, spp = 12, 10, 4
    M = len(sc["mats"])
    cam = scenes.camera_array(sc["camera"])
    rng = np.random.Generator(np.random.PCG64(3))
    dldi = rng.normal(size=(H, W, 3))
    o = po.Scene(sc)
    for tid in (0, 1):
        tex = sc["textures"][tid].copy()
        g, gt = o.adjoint(cam, W, H, spp, dldi, M, seed=13, tex_id=tid, tex_shape=tex.shape[:2])
        assert np.abs(gt).max() > 0
        # the textured material's Kd row gets no gradient
        assert np.all(g[tid, :3] == 0)
        flat = np.argsort(-np.abs(gt).reshape(-1))[:5]
        for f in flat:
            y, x, c = np.unravel_index(f, gt.shape)
            h = 1e-3
            tp, tm = tex.copy(), tex.copy()
            tp[y, x, c] += h
            tm[y, x, c] -= h
            o.set_texture_rgb32f(tid, tp)
            lp = (dldi.reshape(-1) * o.render(cam, W, H, spp, seed=13, fp64=True)[0] / spp).sum()
            o.set_texture_rgb32f(tid, tm)
            lm = (dldi.reshape(-1) * o.render(cam, W, H, spp, seed=13, fp64=True)[0] / spp).sum()
            fd = (lp - lm) / (float(tp[y, x, c]) - float(tm[y, x, c]))
            # the texture lookup is FP32 (texture.cpp), so the difference carries ~1e-7 / h rounding
            assert abs(fd - gt[y, x, c]) <= 1e-4 * abs(fd) + 1e-9, (tid, y, x, c, fd, gt[y, x, c])
        o.set_texture_rgb32f(tid, tex)
