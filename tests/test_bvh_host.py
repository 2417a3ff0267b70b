"""The product's host SAH builder (libnuray, no device needed) reproduces the oracle's tree bit
for bit: same leaves in the same order, same float boxes, same preorder topology."""
import numpy as np
import pytest


def _same(a, b):
    for k in ("left", "right", "leaf_first", "leaf_count", "prim_order"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["box"].view(np.uint32), b["box"].view(np.uint32))


def _dup_scene(scenes):
    """Duplicated geometry: every pMax value is tied, so std::sort's tie order decides the tree."""
    sc = scenes.box_room()
    sc["p"] = np.concatenate([sc["p"]] * 7)
    sc["mat"] = np.concatenate([sc["mat"]] * 7)
    return sc


CASES = {
    "box_room": lambda s: s.box_room(),
    "box_room_64mat": lambda s: s.box_room(patches=True),
    "ties": _dup_scene,
    "terrain64": lambda s: s.terrain_spheres(n_cells=64, n_spheres=100),
    "terrain512": lambda s: s.terrain_spheres(),
}


@pytest.mark.parametrize("name", list(CASES))
def test_product_builder_matches_oracle(name, po, scenes):
    import nuray_amd as nr

    sc = CASES[name](scenes)
    ref = po.Scene(sc).export()
    for threads in (1, 0):
        got = nr.bvh_build_host(sc["p"], nthreads=threads)
        _same(ref, got)


def test_degenerate_inputs(po):
    import nuray_amd as nr

    for n in (1, 5, 6):
        p = np.random.default_rng(n).uniform(0, 1, (n, 9)).astype(np.float32)
        sc = dict(p=p, n=None, uv=None, mat=np.zeros(n, np.int32), mats=np.array([[0, 1, 1, 1, .5, 1, -1, 0]], np.float32))
        _same(po.Scene(sc).export(), nr.bvh_build_host(p))
    # zero-area and identical triangles
    p = np.zeros((9, 9), np.float32)
    sc = dict(p=p, n=None, uv=None, mat=np.zeros(9, np.int32), mats=np.array([[0, 1, 1, 1, .5, 1, -1, 0]], np.float32))
    _same(po.Scene(sc).export(), nr.bvh_build_host(p))


def test_degenerate_split_is_reported_not_a_crash(po):
    """Nested squares growing by 1.15x per step reach 1.5e18: their FP32 surface areas overflow,
    no split has a finite SAH cost, and the reference's recursion would never end.  Both the
    library's builder and the oracle's report it as an error."""
    tris = []
    for k in range(300):
        s, z = 1.15 ** k, -float(k)
        a, b, c, d = [-s, -s, z], [s, -s, z], [s, s, z], [-s, s, z]
        tris += [a + b + c, a + c + d]
    p = np.array(tris, np.float32)
    import nuray_amd as nr

    with pytest.raises(nr.NuRayError, match="finite cost"):
        nr.bvh_build_host(p)
    sc = dict(p=p, n=None, uv=None, mat=np.zeros(len(p), np.int32),
              mats=np.array([[0, .5, .5, .5, .5, 1, -1, 0]], np.float32))
    with pytest.raises(RuntimeError, match="finite cost"):
        po.Scene(sc)
