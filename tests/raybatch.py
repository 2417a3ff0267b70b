"""Seeded ray batches for traversal parity: primary, uniform, secondary and adversarial rays."""
from __future__ import annotations

import numpy as np


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _camera_rays(sc, n, rng):
    cam = sc["camera"]
    pos = np.asarray(cam["pos"], np.float64)
    yaw, pitch, _ = (np.radians(a) for a in cam["euler"])
    gaze = np.array([-np.cos(pitch) * np.sin(yaw), np.sin(pitch), -np.cos(yaw) * np.cos(pitch)])
    up = np.array([np.sin(yaw) * np.sin(pitch), np.cos(pitch), np.cos(yaw) * np.sin(pitch)])
    hand = _unit(np.cross(gaze, up))
    fh = 2 * np.tan(np.radians(cam["fov"]) / 2)
    fw = fh * cam["aspect"]
    cx = rng.uniform(-1, 1, n)[:, None]
    cy = rng.uniform(-1, 1, n)[:, None]
    d = _unit(gaze + hand * fw * cx * 0.5 - up * fh * cy * 0.5)
    return np.repeat(pos[None], n, 0), d


def _bounds(sc):
    p = sc["p"].reshape(-1, 3)
    return p.min(0).astype(np.float64), p.max(0).astype(np.float64)


def adversarial_rays(sc, rng, n):
    """Axis-aligned directions with exact zeros, +-1e-9 components, rays aimed at vertices, edge
    midpoints and box corners, rays lying in triangle planes (grazing) and rays starting on faces."""
    lo, hi = _bounds(sc)
    p = sc["p"].reshape(-1, 3, 3).astype(np.float64)
    out_o, out_d = [], []
    k = max(n // 6, 1)
    # axis-aligned directions from random interior origins
    axes = np.concatenate([np.eye(3), -np.eye(3)])
    o = rng.uniform(lo, hi, (k, 3))
    out_o.append(o); out_d.append(axes[rng.integers(0, 6, k)])
    # nearly axis-aligned: components of +-1e-9 and exact zeros
    d = axes[rng.integers(0, 6, k)].copy()
    tiny = rng.choice([-1e-9, 1e-9, 0.0, -1e-9 * 0.5], (k, 3))
    d = d + tiny * (d == 0)
    out_o.append(rng.uniform(lo, hi, (k, 3))); out_d.append(d)
    # aimed at vertices, edge midpoints and centroids of random triangles
    t = p[rng.integers(0, len(p), k)]
    w = rng.choice(3, k)
    targets = np.where((w == 0)[:, None], t[:, 0], np.where((w == 1)[:, None], 0.5 * (t[:, 0] + t[:, 1]),
                                                             (t[:, 0] + t[:, 1] + t[:, 2]) / 3))
    o = rng.uniform(lo, hi, (k, 3))
    out_o.append(o); out_d.append(_unit(targets - o))
    # grazing: origin in a triangle's plane, direction along an edge
    t = p[rng.integers(0, len(p), k)]
    a = rng.uniform(0, 1, (k, 1))
    o = t[:, 0] + a * (t[:, 1] - t[:, 0]) - 5.0 * _unit(t[:, 2] - t[:, 0])
    out_o.append(o); out_d.append(_unit(t[:, 2] - t[:, 0]))
    # starting exactly on a triangle, leaving along +-normal or the reverse edge
    t = p[rng.integers(0, len(p), k)]
    bc = rng.dirichlet([1, 1, 1], k)
    o = (bc[:, :, None] * t).sum(1)
    nrm = _unit(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]))
    sgn = rng.choice([-1.0, 1.0], (k, 1))
    out_o.append(o); out_d.append(nrm * sgn)
    # aimed at node-box-like corners: origin outside the scene bounds toward its corners
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    c = corners[rng.integers(0, 8, k)]
    o = (lo + hi) / 2 + (hi - lo) * rng.uniform(-1.5, 1.5, (k, 3))
    out_o.append(o); out_d.append(_unit(c - o + rng.choice([0.0, 1e-9], (k, 3))))
    o = np.concatenate(out_o)[:n]; d = np.concatenate(out_d)[:n]
    return o.astype(np.float32), d.astype(np.float32)


def make_batch(sc, n, seed=7, oracle=None):
    """n rays: 1/4 camera, 1/4 uniform (origin in the scene box, direction on the sphere),
    1/4 secondary (cosine-ish bounce from the oracle's first hits, offset 1e-3) and 1/4 adversarial."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = n // 4
    co, cd = _camera_rays(sc, q, rng)
    lo, hi = _bounds(sc)
    uo = rng.uniform(lo, hi, (q, 3))
    ud = _unit(rng.normal(size=(q, 3)))
    so, sd = co[:q], cd[:q]
    if oracle is not None:
        tri, t, _, _ = oracle.intersect(co.astype(np.float32), cd.astype(np.float32))
        ok = tri >= 0
        hp = co + cd * np.where(ok, t, 0.0)[:, None]
        p = sc["p"].reshape(-1, 3, 3).astype(np.float64)
        nrm = _unit(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]))[np.maximum(tri, 0)]
        nrm = np.where((nrm * cd).sum(1, keepdims=True) > 0, -nrm, nrm)
        bd = _unit(nrm + _unit(rng.normal(size=(q, 3))))
        sd = np.where(ok[:, None], bd, ud)
        so = np.where(ok[:, None], hp + sd * 1e-3, uo)
    ao, ad = adversarial_rays(sc, rng, n - 3 * q)
    o = np.concatenate([co, uo, so, ao]).astype(np.float32)
    d = np.concatenate([cd, ud, sd, ad]).astype(np.float32)
    return o, d
