"""Independent pins of four oracle functions that the same-seed GPU tests rely on:

* the GGX-reflect BSDF value with the reference's non-Smith G and Schlick F
  (src/material/matggxrefl.cpp:45-50);
* the translucent GGX model MatGGX: its reflection and transmission values, every pdf branch and
  its sampler, including eta = 1/ior on both sides (src/material/matggx.cpp:26-47,79-105,123-148);
* the direction -> (u, v) lookup of a ray that leaves the scene
  (src/renderer/rendererptls.cpp:24-26, src/scene/texture.cpp:61-72, src/scene/envmap.cpp:17-24);
* the default MIS weight (mis = 1), computed from the NEE direction and carried to the next
  vertex's emitter hit (src/renderer/rendererptls.cpp:61-69,89).

Each expected value comes from a closed form, a published formula or a quadrature written here in
float64, never from the oracle; each test also shows that a plausible mistake (a Smith G1 in the
reflect model, a dropped 0.9999, a 'fixed' eta, a u/v swap or v flip, the textbook MIS weight on
the BSDF direction) would fail it.  CPU only.
"""
import math

import numpy as np
import pytest

PI_REF = 3.14159          # the reference's pi literal (double in some expressions)
PI_REF_F = float(np.float32(3.14159))


def _mat_scene(po, rows):
    p = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    return po.Scene(dict(p=p, n=None, uv=None, mat=np.zeros(1, np.int32), mats=np.array(rows, np.float32)),
                    build_bvh=False)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _dir(theta, phi):
    """Unit vector at polar angle theta from +z, azimuth phi."""
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


# ------------------------------------------------------------------ GGX-reflect value
def _g_ref(c, alpha):
    """The reference's shadowing term G(c) = 2c / (c + sqrt(k^2 + 1 - k^2 c^2)), k = alpha/2 + 1/2."""
    k = alpha * 0.5 + 0.5
    return 2 * c / (c + math.sqrt(k * k + (1 - k * k * c * c)))


def _g_smith(c, alpha):
    """Smith's G1 for the GGX distribution (Walter et al. 2007, eq. 34) at cos theta = c."""
    return 2 * c / (c + math.sqrt(alpha * alpha + (1 - alpha * alpha) * c * c))


def _ndf(c, alpha):
    """GGX normal distribution with the reference's pi literal (matggxrefl.cpp:46)."""
    return alpha * alpha / PI_REF_F / (c * c * (alpha * alpha - 1) + 1) ** 2


@pytest.mark.parametrize("alpha", [0.12, 0.35, 0.8])
def test_ggx_reflect_value_at_the_specular_direction_isolates_g(po, alpha):
    """wi = mirror of wo about n puts the half vector on n: D = 1/(pi' alpha^2) and F = F0 +
    (1 - F0)(1 - cos)^5 are then closed forms, and f * 4 (cos + 1e-6)^2 / (D F) is G(cos)^2 alone.
    It must be the reference's G; Smith's G1 (the textbook choice) differs by far more."""
    F0 = np.array([0.9, 0.5, 0.1])
    s = _mat_scene(po, [[2, *F0, alpha, 1, -1, 0]])
    thetas = np.radians([5.0, 25.0, 45.0, 60.0, 70.0, 80.0])
    n = np.tile([0.0, 0.0, 1.0], (len(thetas), 1))
    wo = np.stack([_dir(t, 0.4) for t in thetas])
    wi = np.stack([_dir(t, 0.4 + math.pi) for t in thetas])
    _, f, pdf = s.material_eval(0, wo, n, wi=wi)
    c = np.cos(thetas)
    D = 1.0 / (PI_REF_F * alpha * alpha)
    for j, cj in enumerate(c):
        F = F0 + (1 - F0) * (1 - cj) ** 5
        want = D * _g_ref(cj, alpha) ** 2 * F / 4 / (cj + 1e-6) ** 2
        assert np.all(_rel(f[j], want) < 5e-5), (thetas[j], f[j], want)
        smith = D * _g_smith(cj, alpha) ** 2 * F / 4 / (cj + 1e-6) ** 2
        if thetas[j] > math.radians(20):
            assert np.all(_rel(f[j], smith) > 1e-2), "a Smith G1 would pass this pin"
        # the pdf at the same pair: D (n.h) / 4 + 1e-8 with n.h = 1, no 1/(4 wo.h) Jacobian
        assert _rel(pdf[j], D / 4 + 1e-8) < 5e-5


def test_ggx_reflect_value_retro_and_generic_pairs(po):
    """Retro-reflection wi = wo puts h on wo: F = F0 exactly, D at n.h = cos.  Generic pairs are
    evaluated from the half-vector angles in float64; the channel ratio f_r / f_b isolates Schlick's
    F at h.wi (a Fresnel term at n.wi fails it); the value is reciprocal in (wo, wi)."""
    alpha = 0.4
    F0 = np.array([0.95, 0.6, 0.05])
    s = _mat_scene(po, [[2, *F0, alpha, 1, -1, 0]])
    # retro-reflection
    thetas = np.radians([10.0, 40.0, 70.0])
    wo = np.stack([_dir(t, 1.1) for t in thetas])
    n = np.tile([0.0, 0.0, 1.0], (len(thetas), 1))
    _, f, _ = s.material_eval(0, wo, n, wi=wo)
    for j, t in enumerate(thetas):
        c = math.cos(t)
        want = _ndf(c, alpha) * _g_ref(c, alpha) ** 2 * F0 / 4 / (c + 1e-6) ** 2
        assert np.all(_rel(f[j], want) < 2e-5)
    # generic pairs about random normals (both directions reflected into the upper hemisphere)
    rng = np.random.Generator(np.random.PCG64(11))
    pairs = []
    while len(pairs) < 12:
        nn = _unit(rng.normal(size=3))
        a, b = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        a, b = a - 2 * min(a @ nn, 0) * nn, b - 2 * min(b @ nn, 0) * nn
        if a @ nn > 0.05 and b @ nn > 0.05:
            pairs.append((nn, a, b))
    N = np.stack([p[0] for p in pairs]); WO = np.stack([p[1] for p in pairs]); WI = np.stack([p[2] for p in pairs])
    _, f, _ = s.material_eval(0, WO, N, wi=WI)
    _, f_swap, _ = s.material_eval(0, WI, N, wi=WO)
    assert np.all(_rel(f_swap, f) < 1e-5)                      # reciprocity
    wrong_fresnel = 0
    for j, (nn, a, b) in enumerate(pairs):
        h = _unit(a + b)
        ch = nn @ h                                            # cos theta_h
        ci, co = nn @ b, nn @ a
        x = (1 - h @ b) ** 5
        want = _ndf(ch, alpha) * _g_ref(ci, alpha) * _g_ref(co, alpha) * (F0 + (1 - F0) * x) / 4 \
            / (co + 1e-6) / (ci + 1e-6)
        assert np.all(_rel(f[j], want) < 5e-5), (j, f[j], want)
        # Schlick's weight read back from two channels: F_r/F_b = (a_r + (1-a_r) x)/(a_b + (1-a_b) x)
        r = float(f[j, 0]) / float(f[j, 2])
        x_back = (r * F0[2] - F0[0]) / ((1 - F0[0]) - r * (1 - F0[2]))
        assert abs(x_back - x) < 1e-5 + 1e-3 * x, (x_back, x)
        wrong_fresnel += abs(x_back - (1 - ci) ** 5) > 1e-3      # a Fresnel weight at n.wi
    assert wrong_fresnel >= len(pairs) // 2
    # below the horizon on one side: zero (matggxrefl.cpp:37-38)
    _, f, _ = s.material_eval(0, WO[:1], N[:1], wi=-WI[:1])
    assert not f.any()


# ------------------------------------------------------------------ MatGGX (translucent)
def _ggx_cdf(c, alpha):
    """Fraction of GGX half vectors (density D(h) n.h) with cos theta_h >= c."""
    return (1 - c * c) / (c * c * (alpha * alpha - 1) + 1)


def test_ggx_theta_inverse_cdf_is_the_ndf_cdf():
    """The closed-form CDF used below equals the quadrature of D(h) cos theta_h over the cap."""
    for alpha in (0.1, 0.5, 0.9):
        th = np.linspace(0, math.pi / 2, 400001)
        c = np.cos(th)
        dens = alpha * alpha / math.pi / (c * c * (alpha * alpha - 1) + 1) ** 2 * c * np.sin(th) * 2 * math.pi
        cum = np.concatenate([[0], np.cumsum((dens[1:] + dens[:-1]) / 2 * np.diff(th))])
        for k in (40000, 150000, 300000):
            assert abs(cum[k] - _ggx_cdf(c[k], alpha)) < 1e-6


def _frame(nrm):
    """Tangent axes of the reference's shading frame (matggx.cpp:22-24)."""
    ax0 = np.array([1.0, 0, 0]) if abs(nrm[0]) < 0.8 else np.array([0, 1.0, 0])
    ax1 = _unit(np.cross(ax0, nrm))
    ax2 = _unit(np.cross(nrm, ax1))
    return ax1, ax2


@pytest.mark.parametrize("alpha", [0.15, 0.5])
def test_matggx_sampler_half_vector_and_eta(po, alpha):
    """MatGGX::sampleBxdf with controlled uniforms.  Reflection (u2 < 0.5): the half vector read back
    from wi = 2 (h.wo) h - wo has theta_h / 0.9999 at the NDF's inverse CDF of u0 and azimuth
    2 pi' u1 in the reference's frame.  Refraction (u2 >= 0.5) with the same (u0, u1): wi lies in
    the plane of wo and h, and its coefficients wi ~ -eta wo - |(eta + 1) wo.h| h give back
    eta = 1/ior for wo on either side (matggx.cpp:43 reads a zero wi, so the outside branch is never
    taken): a 'fixed' eta = ior for wo inside fails."""
    ior = 1.5
    s = _mat_scene(po, [[3, 0.04, 0.04, 0.04, alpha, ior, -1, 0]])
    rng = np.random.Generator(np.random.PCG64(21))
    k = 400
    u = rng.uniform(0.05, 0.95, (k, 3)).astype(np.float32)
    nrm = _unit([0.3, -0.2, 0.93])
    for side in (1.0, -1.0):
        # wo on the outside (side = 1) or inside (side = -1), oblique
        wo = _unit(side * nrm + 0.45 * np.cross(nrm, [0, 0, 1.0]))
        wos = np.tile(wo, (k, 1)); ns = np.tile(nrm, (k, 1))
        ur = u.copy(); ur[:, 2] = 0.25
        wr, _, _ = s.material_eval(0, wos, ns, u3=ur)
        hv = _unit(wr.astype(np.float64) + wo)
        hv *= np.sign(hv @ nrm)[:, None]
        ax1, ax2 = _frame(nrm)
        th = np.arctan2(np.linalg.norm(np.cross(hv, nrm), axis=1), hv @ nrm)
        u0_back = _ggx_cdf(np.cos(th / 0.9999), alpha)
        ok = np.abs(hv @ wo) > 0.05           # h read back reliably
        assert ok.sum() > 300
        assert np.abs(u0_back - u[:, 0])[ok].max() < 6e-6
        u0_nofac = _ggx_cdf(np.cos(th), alpha)   # without the 0.9999 scale every sample is off
        assert np.abs(u0_nofac - u[:, 0])[ok].min() > 6e-6
        phi = np.arctan2(hv @ ax2, hv @ ax1) % (2 * math.pi)
        dphi = (phi - 2 * PI_REF_F * u[:, 1].astype(np.float64) + math.pi) % (2 * math.pi) - math.pi
        assert np.abs(dphi)[ok].max() < 1e-4
        # refraction branch with the same half vectors
        ut = u.copy(); ut[:, 2] = 0.75
        wt, _, _ = s.material_eval(0, wos, ns, u3=ut)
        wt = wt.astype(np.float64)
        etas = []
        for j in np.nonzero(ok)[0]:
            A = np.stack([wo, hv[j]], 1)
            coef, res, _, _ = np.linalg.lstsq(A, wt[j], rcond=None)
            assert np.linalg.norm(A @ coef - wt[j]) < 1e-4       # coplanar with wo and h
            a, b = coef
            r = a / b * abs(wo @ hv[j])
            etas.append(r / (1 - r))
        etas = np.array(etas)
        assert np.abs(etas - 1 / ior).max() < 2e-3, (side, etas.min(), etas.max())
        assert np.abs(etas - ior).min() > 0.5                    # eta = ior on the inside would fail


def _walter_btdf(i, o, n, alpha, ior_in, ior_out, F0):
    """Walter et al. 2007, eq. 21 with eq. 16 (generalised half vector), eq. 33 (GGX D) and eq. 34
    (Smith G1), all with chi+ factors; Fresnel as in the reference: Schlick at (i.h) with F0 and
    zero for i.h < 0 (matggx.cpp:50-55).  i is the direction on the ior_in side, o on ior_out."""
    h = -(ior_in * i + ior_out * o)
    h = h / np.linalg.norm(h)
    if h @ n < 0:
        return np.zeros(3)                                          # D has chi+(m.n)

    def G1(v):
        if (v @ h) * (v @ n) < 0:
            return 0.0
        c = v @ n
        return 2 / (1 + math.sqrt(1 + alpha * alpha * (1 - c * c) / (c * c)))
    c = h @ n
    D = alpha * alpha / PI_REF_F / (c * c * (alpha * alpha - 1) + 1) ** 2
    F = np.zeros(3) if i @ h < 0 else F0 + (1 - F0) * (1 - i @ h) ** 5
    jac = abs((i @ h) * (o @ h) / ((i @ n) * (o @ n)))
    return jac * ior_out ** 2 * (1 - F) * G1(i) * G1(o) * D / (ior_in * (i @ h) + ior_out * (o @ h)) ** 2


def test_matggx_values_reflection_and_transmission(po):
    """MatGGX::bxdf.  Reflection: at normal incidence F = F0, G1 = 1, D = 1/(pi' alpha^2) -> f =
    F0 / (pi' alpha^2) / (4 + 1e-6); at the specular pair f 4 cos^2 / (D F) is Smith's G1^2 (the
    reflect model's G fails); both directions inside give the mirrored value (nn = -n).
    Transmission: Walter's BTDF with eta = eta_o / eta_i (ior when wi is outside, 1/ior inside)
    on refraction pairs that obey Snell's law through a tilted microfacet; swapping the two eta
    conventions fails."""
    alpha, ior = 0.3, 1.5
    F0 = np.array([0.04, 0.2, 0.6])
    s = _mat_scene(po, [[3, *F0, alpha, ior, -1, 0]])
    z = np.array([0.0, 0.0, 1.0])
    _, f, _ = s.material_eval(0, z[None], z[None], wi=z[None])
    assert np.all(_rel(f[0], F0 / (PI_REF_F * alpha * alpha) / (4 + 1e-6)) < 2e-6)
    for t in np.radians([20.0, 50.0, 70.0]):
        wo, wi = _dir(t, 0.3), _dir(t, 0.3 + math.pi)
        _, f, _ = s.material_eval(0, wo[None], z[None], wi=wi[None])
        c = math.cos(t)
        F = F0 + (1 - F0) * (1 - c) ** 5
        D = 1 / (PI_REF_F * alpha * alpha)
        want = F * _g_smith(c, alpha) ** 2 * D / (4 * c * c + 1e-6)
        assert np.all(_rel(f[0], want) < 2e-5)
        assert np.all(_rel(f[0], F * _g_ref(c, alpha) ** 2 * D / (4 * c * c + 1e-6)) > 1e-2)
        _, f_in, _ = s.material_eval(0, -wo[None], z[None], wi=-wi[None])
        assert np.all(_rel(f_in[0], f[0]) < 1e-6)
    # transmission pairs: microfacet m tilted from n, wi outside refracting into wo inside
    rng = np.random.Generator(np.random.PCG64(5))
    checked = 0
    for _ in range(40):
        m = _unit(z + 0.35 * rng.normal(size=3) * np.array([1, 1, 0]))
        i = _unit(rng.normal(size=3))
        if i @ m < 0.2 or i @ z < 0.1:
            continue
        # Snell through m from outside (1) into ior: refracted direction o (pointing away, inside)
        eta = 1 / ior
        ci = i @ m
        ct = math.sqrt(1 - eta * eta * (1 - ci * ci))
        o = -(eta * i) + (eta * ci - ct) * m          # transmitted propagation direction ...
        o = o / np.linalg.norm(o)                      # ... is the direction the light leaves to
        if o @ z > -0.1:
            continue
        for wi_dir, wo_dir in ((i, o), (o, i)):        # light from outside, and from inside
            _, f, _ = s.material_eval(0, wo_dir[None], z[None], wi=wi_dir[None])
            ior_i = 1.0 if wi_dir @ z > 0 else ior
            ior_o = ior if wi_dir @ z > 0 else 1.0
            want = _walter_btdf(wi_dir, wo_dir, z, alpha, ior_i, ior_o, F0)
            assert np.all(want > 0)
            assert np.all(_rel(f[0], want) < 1e-4), (f[0], want)
            wrong = _walter_btdf(wi_dir, wo_dir, z, alpha, ior_o, ior_i, F0)
            assert np.all(_rel(f[0], wrong) > 1e-2)
            checked += 1
    assert checked >= 10


def test_matggx_pdf_branches(po):
    """MatGGX::pdf (matggx.cpp:123-148): same side above -> D (n.h)/4; both below -> 1e18 (D is
    taken about the unflipped n, where the half vector points down); opposite sides -> half vector
    -(wi + eta wo) with eta = ior if wi is outside else 1/ior, flipped up, D (n.h)/4 clamped below
    at 1, and 1e18 when D < 1e-6."""
    ior = 1.5
    z = np.array([0.0, 0.0, 1.0])
    s = _mat_scene(po, [[3, .04, .04, .04, 0.3, ior, -1, 0], [3, .04, .04, .04, 0.05, ior, -1, 0],
                        [3, .04, .04, .04, 1e-3, ior, -1, 0]])
    wo, wi = _dir(0.5, 0.2), _dir(0.9, 2.0)
    _, _, p = s.material_eval(0, wo[None], z[None], wi=wi[None])
    h = _unit(wo + wi)
    assert _rel(p[0], _ndf(h @ z, 0.3) * (h @ z) / 4) < 2e-6
    _, _, p = s.material_eval(0, -wo[None], z[None], wi=-wi[None])
    assert p[0] == np.float32(1e18)
    # physical refraction pair through m = n: D at its peak
    for a_idx, alpha in ((0, 0.3), (1, 0.05)):
        sin_i = math.sin(0.6)
        wi_out = _dir(0.6, 0.0)                                   # outside
        wo_in = _dir(math.pi - math.asin(sin_i / ior), math.pi)  # inside, Snell through n
        _, _, p = s.material_eval(a_idx, wo_in[None], z[None], wi=wi_out[None])
        peak = 1 / (PI_REF_F * alpha * alpha) / 4
        assert _rel(p[0], max(peak, 1.0)) < 1e-4, (alpha, p[0], peak)
        _, _, p2 = s.material_eval(a_idx, wi_out[None], z[None], wi=wo_in[None])   # light from inside
        assert _rel(p2[0], max(peak, 1.0)) < 1e-4
    # alpha = 0.05 has peak 31.8 > 1; with the eta convention swapped the half vector tilts far from
    # n and D collapses, so this pair separates the two conventions
    hw = -(wi_out + (1 / ior) * wo_in); hw = _unit(hw); hw *= np.sign(hw @ z)
    wrong = max(_ndf(hw @ z, 0.05) * (hw @ z) / 4, 1.0)
    assert abs(wrong - 1 / (PI_REF_F * 0.05 ** 2) / 4) > 10
    # D < 1e-6 for alpha = 1e-3 away from the peak: 1e18
    # (wi + ior wo has almost no z component, so the half vector is nearly horizontal)
    _, _, p = s.material_eval(2, _dir(math.pi - 1.2, 0.0)[None], z[None], wi=_dir(1.0, 0.0)[None])
    assert p[0] == np.float32(1e18)
    _, _, p = s.material_eval(1, _dir(math.pi - 1.2, 0.0)[None], z[None], wi=_dir(1.0, 0.0)[None])
    assert p[0] == np.float32(1.0)                                # alpha = 0.05: D = 8e-4, clamped


# ------------------------------------------------------------------ miss direction -> envmap
def _env_scene(tex, gain=1.0):
    R = 1.0
    p = np.array([[-R, -1e3, R, R, -1e3, R, R, -1e3, -R]], np.float32)
    return dict(p=p, n=None, uv=None, mat=np.zeros(1, np.int32),
                mats=np.array([[0, .5, .5, .5, .5, 1, -1, 0]], np.float32), envmap=dict(rgb=tex, gain=gain))


def _env_texels():
    """4 x 2 sky with a distinct value per texel and channel."""
    H, W = 2, 4
    t = np.zeros((H, W, 3), np.float32)
    for i in range(H):
        for j in range(W):
            k = 1 + j + W * i
            t[i, j] = (k, 10.0 * k, 100.0 * k)
    return t


def _bilinear_black(t, x, y):
    """texture.cpp:54-67 in float64: corners floor(x), floor(x)+1, black outside the image."""
    H, W = t.shape[:2]
    x0, y0 = math.floor(x), math.floor(y)
    out = np.zeros(3)
    for xx, wx in ((x0, x0 + 1 - x), (x0 + 1, x - x0)):
        for yy, wy in ((y0, y0 + 1 - y), (y0 + 1, y - y0)):
            if 0 <= xx < W and 0 <= yy < H:
                out += wx * wy * t[yy, xx].astype(np.float64)
    return out


def _aim(x, y, W, H):
    """Direction whose lookup lands on texel coordinate (x, y): 1 - v = acos(d.y)/pi' = y/H and
    u = atan2(d.z, d.x)/(2 pi') + 1/2 = x/W (rendererptls.cpp:24-25)."""
    th = PI_REF * y / H
    ph = (x / W - 0.5) * 2 * PI_REF
    return np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])


def test_miss_direction_uv_mapping(po):
    """Rays aimed at (j + 1/4, i + 1/4) of every texel of a 4 x 2 sky return the bilinear blend
    9/16 T(j,i) + 3/16 T(j+1,i) + 3/16 T(j,i+1) + 1/16 T(j+1,i+1) (black past the last column and
    row) times the gain.  A u/v swap, a flipped v or a flipped u would return other texels."""
    t = _env_texels()
    H, W = t.shape[:2]
    gain = 1.5
    o = po.Scene(_env_scene(t, gain), build_bvh=False)
    pts = [(j + 0.25, i + 0.25) for i in range(H) for j in range(W)] + [(j + 0.7, i + 0.6) for i in range(H)
                                                                       for j in range(W)]
    d = np.stack([_aim(x, y, W, H) for x, y in pts])
    got = o.env_radiance(d).astype(np.float64)
    want = np.stack([_bilinear_black(t, x, y) for x, y in pts]) * gain
    assert np.all(_rel(got, want) < 1e-4), np.max(_rel(got, want))
    # the plausible slips return other values at (almost) every aimed point
    slips = {
        "u/v swapped": [_bilinear_black(t, y / H * W, x / W * H) for x, y in pts],
        "v flipped": [_bilinear_black(t, x, (H - y) % H) for x, y in pts],
        "u flipped": [_bilinear_black(t, (W - x) % W, y) for x, y in pts],
    }
    for name, vals in slips.items():
        bad = np.stack(vals) * gain
        assert (np.max(_rel(got, bad), axis=1) > 1e-2).mean() > 0.8, name
    # the seam: u just below 1 blends the last column with the black border (texture.cpp:54-59)
    g = o.env_radiance(_aim(W - 0.5, 0.25, W, H)[None])[0]
    assert np.all(_rel(g, 0.5 * 0.75 * t[0, W - 1] * gain + 0.5 * 0.25 * t[1, W - 1] * gain) < 1e-4)


def test_miss_radiance_through_the_estimator(po):
    """A camera ray that leaves the scene returns the sky at its direction with no MIS weight
    (rendererptls.cpp:16-28): a 1-pixel camera with a 0.01 degree field aimed at texel points
    sees the same blend as the direct lookup."""
    t = _env_texels()
    H, W = t.shape[:2]
    o = po.Scene(_env_scene(t, 1.0))
    for x, y in [(0.25, 0.4), (1.6, 0.7), (2.3, 1.2), (3.25, 1.3)]:
        d = _aim(x, y, W, H)
        pitch = math.degrees(math.asin(d[1]))
        yaw = math.degrees(math.atan2(-d[0], -d[2]))       # gaze = (-cos p sin y, sin p, -cos p cos y)
        cam = np.array([0, 0, 0, yaw, pitch, 0, 0.01, 1], np.float32)
        s, _ = o.render(cam, 1, 1, 16, seed=1)
        got = s.reshape(3) / 16
        want = _bilinear_black(t, x, y)
        assert np.all(_rel(got, want) < 2e-3), (x, y, got, want)


# ------------------------------------------------------------------ default MIS weight (mis = 1)
def _light_plane(rho, Le, a, h):
    R = 1e4
    p = np.array([[-R, 0, R, R, 0, R, R, 0, -R], [R, 0, -R, -R, 0, -R, -R, 0, R],
                  [-a / 2, h, -a / 2, a / 2, h, -a / 2, a / 2, h, a / 2],
                  [a / 2, h, a / 2, -a / 2, h, a / 2, -a / 2, h, -a / 2]], np.float32)
    for k in (2, 3):   # the emitter faces down
        e = np.cross(p[k, 3:6] - p[k, :3], p[k, 6:] - p[k, :3])
        assert e[1] < 0
    return dict(p=p, n=None, uv=None, mat=np.array([0, 0, 1, 1], np.int32),
                mats=np.array([[0, rho, rho, rho, 0.5, 1, -1, 0], [1, Le, Le, Le, 0.5, 1, -1, 0]], np.float32))


def _mis_expectations(rho, Le, a, h, m=1601):
    """Expected radiance at the plane point under the light centre, seen from above, for maxdepth 2
    (vertex 0: NEE and a BSDF-sampled ray; vertex 1: the emitter hit or the black sky), by midpoint
    quadrature over the light square.

    NEE: the light point y is uniform on the square; the reported density is p() = 1/sum area() =
    1/(2 a^2) (area() is twice the triangle area); w_l = p() r^2 / max(1e-6, cos_l), w_b =
    (cos_x + 1e-4)/pi', mis = w_l / (w_l + w_b); the term is Le cos_l / r^2 / p() * rho/pi' * cos_x *
    mis, so E = (2 rho/pi') Le integral_A G mis dA (G = cos_l cos_x / r^2).
    Emitter hit: the BSDF ray (density cos/(0.99 pi) above cos = 0.1, f/pdf = rho/(cos + 1e-4))
    reaches the light with weight fac = 1 - mis(y_NEE) of the NEE sample drawn at the same vertex,
    which is independent of the ray: E = (1 - E_A[mis]) rho Le integral cos^2/(0.99 pi (cos + 1e-4))
    dw over the light.  The textbook rule weights the hit by 1 - mis evaluated along the ray itself."""
    xs = (np.arange(m) + 0.5) / m * a - a / 2
    X, Z = np.meshgrid(xs, xs)
    dA = (a / m) ** 2
    r2 = X * X + Z * Z + h * h
    r = np.sqrt(r2)
    cos_x = h / r                      # at the plane point (normal +y)
    cos_l = h / r                      # at the light (normal -y)
    G = cos_x * cos_l / r2
    lp = 1.0 / (2 * a * a)
    wl = lp * r2 / np.maximum(1e-6, cos_l)
    wb = (cos_x + 1e-4) / PI_REF
    mis = wl / (wl + wb)
    e_nee = 2 * rho / PI_REF_F * Le * (G * mis).sum() * dA
    dw = cos_l / r2 * dA                # solid angle of each cell seen from the plane point
    g_hit = cos_x * cos_x / (0.99 * math.pi * (cos_x + 1e-4))
    e_hit_ref = (1 - mis.mean()) * rho * Le * (g_hit * dw).sum()
    e_hit_textbook = rho * Le * (g_hit * (1 - mis) * dw).sum()
    return e_nee, e_hit_ref, e_hit_textbook


def test_default_mis_weight_from_the_nee_direction(po):
    """mis = 1 (rendererptls.cpp:61-69,89): a 10 x 10 light 4 units above a Lambert plane (so w_l and
    w_b cross over on the light); the oracle's mean at 400k samples must equal the quadrature of
    the reference rule within 5 standard errors, and the textbook balance heuristic on the BSDF
    direction must lie more than 5 standard errors away."""
    rho, Le, a, h = 0.5, 4.0, 10.0, 4.0
    sc = _light_plane(rho, Le, a, h)
    o = po.Scene(sc)
    cam = np.array([0, h / 2, 0, 0, -90, 0, 0.01, 1], np.float32)
    spp = 400000
    s, s2 = o.render(cam, 1, 1, spp, seed=6, mis=1, maxdepth=2)
    m = s[0] / spp
    se = math.sqrt(max(s2[0] / spp - m * m, 0)) / math.sqrt(spp)
    e_nee, e_hit, e_text = _mis_expectations(rho, Le, a, h)
    expect, textbook = e_nee + e_hit, e_nee + e_text
    assert abs(m - expect) < 5 * se + 2e-4 * expect, (m, expect, se)
    assert abs(textbook - expect) > 10 * se, (textbook, expect, se)
    assert abs(m - textbook) > 5 * se
    # the weights really vary over the light (the test would be blind if mis were ~constant)
    assert 0.05 < e_hit / expect < 0.6
