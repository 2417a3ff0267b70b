"""Traversal parity on the GPU: libnuray's nr_intersect against the oracle's restatement of
BVH::intersection (bvh.cpp:99-137) on identical ray batches.

Bar: hit triangle index identical, t / b1 / b2 bit-identical (the required tolerance for t is
1e-5 relative; it is checked too, but the kernel reproduces the reference arithmetic exactly)."""
import numpy as np
import pytest

from raybatch import make_batch

pytestmark = pytest.mark.gpu

SCENES = {
    "box_room": lambda s: s.box_room(),
    "box_room_64mat": lambda s: s.box_room(patches=True),
    "terrain64": lambda s: s.terrain_spheres(n_cells=64, n_spheres=100, width=160, height=90),
    "terrain512": lambda s: s.terrain_spheres(),
}
N_RAYS = {"box_room": 1 << 18, "box_room_64mat": 1 << 18, "terrain64": 1 << 18, "terrain512": 1 << 20}

_cache = {}


def _pair(name, nr, po, scenes):
    if name not in _cache:
        sc = SCENES[name](scenes)
        _cache[name] = (sc, po.Scene(sc), nr.Scene(sc))
    return _cache[name]


def _check_equal(ref, got, label):
    tri_r, t_r, b1_r, b2_r = ref
    tri_g, t_g, b1_g, b2_g = got
    bad = np.nonzero(tri_r != tri_g)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} triangle-index mismatches, first {bad[:5]}: ref {tri_r[bad[:5]]} got {tri_g[bad[:5]]}"
    hit = tri_r >= 0
    rel = np.abs(t_g[hit].astype(np.float64) - t_r[hit]) / np.maximum(np.abs(t_r[hit]), 1e-30)
    assert (rel <= 1e-5).all(), f"{label}: t beyond 1e-5 relative"
    for a, b, nm in ((t_r, t_g, "t"), (b1_r, b1_g, "b1"), (b2_r, b2_g, "b2")):
        nb = np.count_nonzero(a.view(np.uint32) != b.view(np.uint32))
        assert nb == 0, f"{label}: {nb} {nm} values not bit-identical"


@pytest.mark.parametrize("name", list(SCENES))
def test_closest_hit_bit_exact(name, nr, po, scenes):
    sc, ref, dev = _pair(name, nr, po, scenes)
    o, d = make_batch(sc, N_RAYS[name], seed=7, oracle=ref)
    _check_equal(ref.intersect(o, d), dev.intersect(o, d), name)


@pytest.mark.parametrize("name", list(SCENES))
def test_shadow_any_hit_matches_closest_rule(name, nr, po, scenes):
    """Shadow rays: blocked iff the reference's closest hit has (double)t + 2e-3 <= length
    (rendererptls.cpp:70-71).  Lengths straddle the closest hit to probe the boundary."""
    sc, ref, dev = _pair(name, nr, po, scenes)
    o, d = make_batch(sc, min(N_RAYS[name], 1 << 18), seed=9, oracle=ref)
    tri_r, t_r, _, _ = ref.intersect(o, d)
    rng = np.random.Generator(np.random.PCG64(11))
    base = np.where(tri_r >= 0, t_r, rng.uniform(1, 500, len(o))).astype(np.float64)
    jitter = rng.choice([-1e-2, -2e-3, -1e-6, 0.0, 1e-6, 2e-3, 1e-2, 50.0], len(o))
    length = (base + jitter).astype(np.float32)
    exp_blocked = (tri_r >= 0) & (t_r.astype(np.float64) + 2e-3 <= length.astype(np.float64))
    tri_g, t_g, _, _ = dev.intersect(o, d, any_hit=True, tmax=length)
    got_blocked = tri_g >= 0
    assert np.array_equal(got_blocked, exp_blocked), f"{np.count_nonzero(got_blocked != exp_blocked)} mismatches"
    # the reported blocker must itself satisfy the test
    assert (t_g[got_blocked].astype(np.float64) + 2e-3 <= length[got_blocked]).all()


def test_empty_and_tiny_batches(nr, po, scenes):
    sc, ref, dev = _pair("box_room", nr, po, scenes)
    tri, t, _, _ = dev.intersect(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert len(tri) == 0
    o = np.array([[0, 100, 80]], np.float32)
    d = np.array([[0, -1, 0]], np.float32)
    _check_equal(ref.intersect(o, d), dev.intersect(o, d), "single ray")
    tri, t, _, _ = dev.intersect(o, d)
    assert abs(t[0] - 100.0) < 1e-3  # the floor at y = 0


@pytest.mark.parametrize("entries", [1, 2, 4])
def test_short_stack_retry_is_bit_exact(entries, nr, po, scenes):
    """With a first-launch stack of only `entries` LDS entries most rays overflow it and are traced
    again with the full-depth stack: hits stay bit-identical to the oracle, and a whole render
    (closest-hit and shadow launches) equals the default render bit for bit."""
    sc = SCENES["terrain64"](scenes)
    ref = po.Scene(sc)
    dev = nr.Scene(sc)
    dev.set_short_stack(entries)
    o, d = make_batch(sc, 1 << 17, seed=13, oracle=ref)
    _check_equal(ref.intersect(o, d), dev.intersect(o, d), f"short stack {entries}")
    dev.set_profiling(2)
    a = dev.render(sc["camera"], 96, 54, 4, seed=5, maxdepth=6, rrdepth=2)
    retried = dev.stats()["retried"]
    dev.set_short_stack(16)
    b = dev.render(sc["camera"], 96, 54, 4, seed=5, maxdepth=6, rrdepth=2)
    dev.close()
    print(f"short stack {entries}: {retried:.0f} rays retried")
    assert retried > 0
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sum2"], b["sum2"])


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_trace_order_is_bit_exact(mode, nr, scenes):
    """Secondary and shadow rays (modes 1, 2), or shadow rays alone (3, 4), traced in
    direction-sorted order give the queue-order render bit for bit (the winner among equal t is
    fixed by leaf order, not by visit order), also when most rays overflow a 1-entry first-launch
    stack and go through the retry list."""
    sc = SCENES["terrain64"](scenes)
    dev = nr.Scene(sc)
    out = []
    for order, stack in ((0, 16), (mode, 16), (mode, 1)):
        dev.set_trace_order(order)
        dev.set_short_stack(stack)
        out.append(dev.render(sc["camera"], 96, 54, 4, seed=5, maxdepth=6, rrdepth=2))
    with pytest.raises(RuntimeError):
        dev.set_trace_order(5)
    dev.close()
    print(f"trace order {mode}: {len(out)} renders")
    for r in out[1:]:
        assert np.array_equal(r["sum"], out[0]["sum"]) and np.array_equal(r["sum2"], out[0]["sum2"])


def test_deep_chain_tree(nr, po, scenes):
    """100 coincident triangles stacked in the room make a 100-level chain in the reference's SAH
    tree (all splits cost the same, the first one wins).  Rays that reach the stack overflow the
    short LDS stack and are traced again with the full-depth stack of the 4-wide walk (3 entries
    per level); hits stay bit-identical to the oracle, including the tie among the coincident
    triangles (the later leaf in DFS order wins).  The per-lane binary walk of the path kernel
    renders the same image as the wavefront."""
    sc = scenes.box_room()
    k = 100
    tri = np.array([[-20, 60, -10, 20, 60, -10, 0, 60, 30]], np.float32)
    sc["p"] = np.concatenate([sc["p"], np.tile(tri, (k, 1))])
    sc["mat"] = np.concatenate([sc["mat"], np.zeros(k, sc["mat"].dtype)])
    ref = po.Scene(sc)
    dev = nr.Scene(sc)
    assert dev.info()["depth"] >= 100
    rng = np.random.Generator(np.random.PCG64(21))
    o, d = make_batch(sc, 1 << 16, seed=17, oracle=ref)
    n = 1 << 14
    down_o = np.stack([rng.uniform(-25, 25, n), np.full(n, 150.0), rng.uniform(-15, 35, n)], 1)
    down_d = np.tile([[0.0, -1.0, 0.0]], (n, 1))
    graze_o = np.stack([rng.uniform(-25, 25, n), np.full(n, 60.0), np.full(n, -90.0)], 1)
    graze_d = np.tile([[0.0, 0.0, 1.0]], (n, 1))
    o = np.concatenate([o, down_o, graze_o]).astype(np.float32)
    d = np.concatenate([d, down_d, graze_d]).astype(np.float32)
    got = dev.intersect(o, d)
    _check_equal(ref.intersect(o, d), got, "deep chain")
    assert (got[0][len(o) - 2 * n:len(o) - n] >= 36).mean() > 0.2  # the stack covers ~32 % of their square
    dev.set_profiling(2)
    a = dev.render(sc["camera"], 64, 48, 4, seed=3)
    retried = dev.stats()["retried"]
    dev.set_profiling(0)
    b = dev.render(sc["camera"], 64, 48, 4, seed=3, path_kernel=True)
    dev.close()
    print(f"deep chain: {retried:.0f} rays retried")
    assert retried > 0
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sum2"], b["sum2"])
