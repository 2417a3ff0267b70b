"""The device's radiance-cache frame (nr_render_nrc) against the oracle's restatement of the same
frame (oracle/orc_nrc_frame.cpp, pinned in tests/test_oracle_pins_nrc_frame.py) on the same random
numbers and the same network: the image of camera paths that stop at the spread rule and take the
cache's answer (renderernrc.cpp:185-281), the number of cache queries, and the training records with
their targets (L_path - L_before) / beta (renderernrc.cpp:128-133, 175-182), path by path.

Same-seed agreement is per path, not bit for bit: where the FP32 acos / atan2 / pow of the device
and of the host's libm differ by an ulp a path can branch differently (the ptnee same-seed tests
see the same few per cent), and the network sees the query's angles rounded to BF16."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _by_path(recs, tgt):
    """{slot: (records, targets)} with each path's records in vertex order."""
    slots = recs[:, 11].view(np.int32)
    order = np.argsort(slots, kind="stable")
    slots, recs, tgt = slots[order], recs[order], tgt[order]
    starts = np.r_[0, np.flatnonzero(np.diff(slots)) + 1, len(slots)]
    return {int(slots[a]): (recs[a:b], tgt[a:b]) for a, b in zip(starts[:-1], starts[1:])}


@pytest.mark.parametrize("which", ["box_room", "material_zoo"])
def test_cache_frame_matches_oracle(nr, po, scenes, which):
    sc = getattr(scenes, which)()
    W, H, spp, T, maxlen = 64, 48, 2, 8192, 5
    cam = dict(sc["camera"], aspect=W / H)
    dev = nr.Scene(sc)
    cache = nr.Nrc(seed=3)
    w = cache.weights()
    print(f"{which}: device cache frame {W}x{H}x{spp} + {T} training paths", flush=True)
    st, img = dev.render_nrc(cache, cam, W, H, spp=spp, seed=5, train_paths=T, train_maxlen=maxlen, spread=0.01,
                             use_cache=True, train=True, min_depth=1, fetch=True)
    g_recs, g_tgt = dev.nrc_frame_records()
    dev.close()
    o = po.Scene(sc).nrc_frame(scenes.camera_array(cam), W, H, spp, w, seed=5, train_paths=T, train_maxlen=maxlen,
                               c_spread=0.01, use_cache=True, min_depth=1)
    # (1) the image: camera paths, their cache answers
    m_g, m_o = img["sum"].reshape(-1, 3) / spp, o["sum"].reshape(-1, 3) / spp
    off = float((np.abs(m_g - m_o) / (np.abs(m_o) + 1e-3) > 1e-3).any(-1).mean())
    print(f"{which}: image {off:.2%} of pixels beyond 1e-3, means {m_g.mean():.5f} / {m_o.mean():.5f}; queries "
          f"{st['queries']} / {o['queries']}; records {len(g_recs)} / {len(o['recs'])}")
    assert o["queries"] > 0.1 * W * H * spp
    assert off < 0.04 and abs(m_g.mean() - m_o.mean()) < 3e-3 * m_o.mean()
    assert abs(st["queries"] - o["queries"]) <= 0.01 * o["queries"] + 5
    # (2) the training records and targets, path by path
    assert abs(len(g_recs) - len(o["recs"])) <= 0.01 * len(o["recs"]) + 5
    gp, op = _by_path(g_recs, g_tgt[:, :3]), _by_path(o["recs"], o["tgt"])
    match = 0
    worst_feat = worst_tgt = 0.0
    for slot, (ro, to) in op.items():
        if slot not in gp or len(gp[slot][0]) != len(ro):
            continue
        rg, tg = gp[slot]
        feat = np.abs(rg[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14]] -
                      ro[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14]]) / \
            (np.abs(ro[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14]]) + 1e-3)
        dt = np.abs(tg - to) / (np.abs(to) + 1e-2)
        if feat.max() < 1e-4 and dt.max() < 1e-3:
            match += 1
            worst_feat, worst_tgt = max(worst_feat, feat.max()), max(worst_tgt, dt.max())
    frac = match / len(op)
    print(f"{which}: {match} of {len(op)} training paths with every record and target equal (largest relative "
          f"differences {worst_feat:.1e} / {worst_tgt:.1e})")
    assert frac > 0.96
