"""Radiance-cache network on the GPU (BF16 MFMA, FP32 accumulation) against the FP64 oracle
network: exact routing through an identity network, inference and loss gradient against the
oracle with the same BF16 rounding points (and the exact network, stated), batch-size edges,
and training that fits a smooth field."""
import numpy as np
import pytest

import nrcdata

pytestmark = pytest.mark.gpu


def test_identity_network_routes_albedo_exactly(nr):
    """W^T rows are output units, the lane layout of every MFMA operand is right: the BF16
    network with 0/1 weights returns the BF16-rounded albedo bit for bit."""
    net = nr.Nrc(seed=1)
    net.set_weights(nrcdata.identity_weights(nr.NRC_NPARAM))
    q = nrcdata.queries(1000, seed=1)
    y = net.infer(q)
    a = q[:, 8:11].astype(np.float32)
    ab = ((a.view(np.uint32) + 0x7FFF + ((a.view(np.uint32) >> 16) & 1)) & 0xFFFF0000).view(np.float32)
    assert np.array_equal(y, ab)


def test_inference_matches_oracle(nr, po):
    net = nr.Nrc(seed=7)
    w = net.weights()
    q = nrcdata.queries(5000, seed=2)
    y = net.infer(q).astype(np.float64)
    yb = po.nrc_forward(w, q, bf16=True)
    ye = po.nrc_forward(w, q, bf16=False)
    scale = np.abs(yb).max()
    err_b = np.abs(y - yb).max() / scale
    err_e = np.abs(y - ye).max() / scale
    print(f"inference: max |gpu - oracle(bf16 points)| / max|y| = {err_b:.2e}; vs exact FP64 network {err_e:.2e}")
    assert err_b < 5e-3
    assert np.median(np.abs(y - yb)) < 1e-3 * scale
    assert err_e < 5e-2


def test_gradient_matches_oracle(nr, po):
    net = nr.Nrc(seed=11)
    w = net.weights()
    q = nrcdata.queries(3000, seed=3)
    t = nrcdata.smooth_target(q)
    loss, g = net.gradient(q, t)
    lb, gb = po.nrc_grad(w, q, t, bf16=True)
    le, ge = po.nrc_grad(w, q, t, bf16=False)
    rel_b = np.linalg.norm(g - gb) / np.linalg.norm(gb)
    rel_e = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    print(f"gradient: relative L2 vs oracle(bf16 points) {rel_b:.2e}, vs exact FP64 {rel_e:.2e}; "
          f"loss {loss:.6f} / {lb:.6f} / {le:.6f}")
    assert abs(loss - lb) < 2e-3 * lb
    assert rel_b < 1e-2
    assert rel_e < 5e-2
    for l in range(6):   # every layer's block agrees, not only the dominant ones
        sl = slice(l * 4096, l * 4096 + (4096 if l < 5 else 1024))
        assert np.linalg.norm(g[sl] - gb[sl]) <= 2e-2 * np.linalg.norm(gb[sl]) + 1e-12, l
    pad = np.zeros(nr.NRC_NPARAM, bool)
    pad[5 * 4096 + 3 * 64: 5 * 4096 + 1024] = True     # unused output rows
    for o in range(64):
        pad[o * 64 + 45: o * 64 + 64] = True            # layer-0 inputs that are always zero
    assert not g[pad].any()


@pytest.mark.parametrize("n", [1, 63, 65, 1000, 4096 + 17])
def test_batch_edges(nr, po, n):
    net = nr.Nrc(seed=3)
    w = net.weights()
    q = nrcdata.queries(n, seed=n)
    y = net.infer(q).astype(np.float64)
    yb = po.nrc_forward(w, q, bf16=True)
    assert np.abs(y - yb).max() <= 5e-3 * max(np.abs(yb).max(), 1e-3)
    t = nrcdata.smooth_target(q)
    _, g = net.gradient(q, t)
    _, gb = po.nrc_grad(w, q, t, bf16=True)
    assert np.linalg.norm(g - gb) <= 2e-2 * np.linalg.norm(gb)


def test_training_fits_smooth_field(nr):
    net = nr.Nrc(seed=5, lr=1e-3)
    q = nrcdata.queries(1 << 16, seed=9)
    t = nrcdata.smooth_target(q)
    qv = nrcdata.queries(8192, seed=10)
    tv = nrcdata.smooth_target(qv)

    def val():
        y = net.infer(qv)
        return float((((y - tv) ** 2) / (y.astype(np.float64) ** 2 + 0.01)).sum(1).mean())
    v0 = val()
    losses = [net.train_step(q, t) for _ in range(300)]
    v1 = val()
    print(f"training: loss {losses[0]:.4f} -> {losses[-1]:.4f}; held-out {v0:.4f} -> {v1:.4f}")
    assert losses[-1] < 0.1 * losses[0]
    assert v1 < 0.1 * v0


def test_batch_larger_than_max_is_an_error(nr):
    net = nr.Nrc(seed=1, max_batch=256)
    q = nrcdata.queries(300, seed=1)
    with pytest.raises(nr.NuRayError):
        net.train_step(q, nrcdata.smooth_target(q))
