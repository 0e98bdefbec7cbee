"""GPU parity of the warm-start network's bf16 MFMA kernels against the CPU oracle
(oracle/surrogate_oracle.py), through the C-ABI of include/dsmc_surrogate.h.

Layer by layer, the oracle is fed the GPU's own inputs of that layer, so each kernel is checked
on its own: a bf16 result must equal the oracle's rounding of the exact value, except where
FP32 accumulation order moves a value across a rounding boundary (then it is one bf16 step
away, and such elements are rare). Weight and bias gradients are FP32 sums checked to 1e-4.
End to end, the rounded oracle network agrees to the drift those rare steps cause.

Each check runs at a small ragged raster (32 x 24: partial pixel tiles at every level) and at the
network's real shape (256 x 256, C = 32, the raster the warm start predicts on).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import surrogate_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu

C = 32
# (height, width, batch): non-square with ragged pixel tiles at every level, and the real raster
SHAPES = [(32, 24, 2), (256, 256, 2)]


@pytest.fixture(scope="module", params=SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def net(request):
    from dsmc_erlend_amd.surrogate import Surrogate
    H, W, B = request.param
    n = Surrogate(height=H, width=W, max_batch=B, seed=7)
    n.test_shape = (H, W, B)
    yield n
    n.close()


def _shape(net, l):
    H, W, B = net.test_shape
    defs = so.layer_defs(C)
    if l < 14:
        lev, ch = defs[l][4], defs[l][3]
    elif l < 17:
        lev, ch = l - 13, C << (l - 14)
    else:
        lev, ch = defs[l - 17][4], defs[l - 17][3]
    return (B, H >> lev, W >> lev, ch)


def _fetch(net, l):
    return net.activation(l).astype(np.float64).reshape(_shape(net, l))


def _inputs(net, seed):
    H, W, B = net.test_shape
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, H, W, 32))
    x[..., 12:] = 0.0
    x[..., 0] = rng.uniform(size=(B, H, W)) < 0.1
    return so.bf16(x)


def _layer_inputs(g, l):
    src = {0: ("in", None), 2: (14, None), 4: (15, None), 6: (16, None),
           8: (5, 7), 10: (3, 9), 12: (1, 11)}
    k1, k2 = src.get(l, (l - 1, None))
    return g[k1], (g[k2] if k2 is not None else None)


def _assert_bf16_close(got, want, what, steps=1, frac=0.01, addends=None):
    """Equal, or within `steps` bf16 rounding steps on a small fraction of elements. Where
    `want` is a sum, `addends` (the sum of its terms' magnitudes) sets the scale of a rounding
    step: a rounding of the terms moves the sum by a step of the terms, however much they
    cancel."""
    diff = np.abs(got - want)
    scale = np.maximum(np.abs(got), np.abs(want))
    if addends is not None:
        scale = np.maximum(scale, addends)
    atol = 1e-6 * max(np.abs(want).max(), 1e-30)
    ok = diff <= steps * 2.0 ** -7 * scale + atol
    assert ok.all(), f"{what}: {np.count_nonzero(~ok)} elements beyond {steps} bf16 steps, " \
                     f"max diff {diff.max():.3e}"
    moved = np.count_nonzero(got != want) / got.size
    assert moved <= frac, f"{what}: {moved:.4f} of elements differ"


def test_encode_matches_oracle(net):
    from dsmc_erlend_amd import default_params
    cases = [default_params(attack_angle_deg=-22.0, mach=6.5, wall_model=2, acc_n=0.6, acc_t=0.9,
                            n_free=3e16),
             default_params(attack_angle_deg=10.0, mach=2.5, wall_model=1, n_free=8e17)]
    H, W, _ = net.test_shape
    got = net.encode(cases)
    for i, p in enumerate(cases):
        np.testing.assert_array_equal(got[i], so.encode(p, H, W))


def test_forward_layer_by_layer_and_end_to_end(net):
    flat = net.weights()
    x = _inputs(net, 1)
    out = net.forward_raw(x)
    orc = so.Net(flat, C, rounding=True)
    g = {"in": x}
    for l in range(17):
        g[l] = _fetch(net, l)
    for l in range(14):
        x1, x2 = _layer_inputs(g, l)
        _assert_bf16_close(g[l], orc.conv_layer(l, x1, x2), f"conv {so.layer_defs(C)[l][0]}")
    for k, src in ((14, 1), (15, 3), (16, 5)):
        np.testing.assert_array_equal(g[k], orc.pool(g[src]))
    head = g[13] @ orc.wh.T + orc.bh
    np.testing.assert_allclose(out, head, rtol=1e-5, atol=1e-5 * np.abs(head).max())
    full = orc.forward(x)
    rel = np.linalg.norm(out - full) / np.linalg.norm(full)
    assert rel < 1e-2, rel


def test_gradients_layer_by_layer_and_end_to_end(net):
    H, W, B = net.test_shape
    flat = net.weights()
    rng = np.random.default_rng(2)
    x = _inputs(net, 3)
    t = rng.normal(size=(B, H, W, 4))
    w = rng.uniform(size=(B, H, W)) * (rng.uniform(size=(B, H, W)) > 0.3) / (H * W)
    loss, grad = net.loss_grad_raw(x, t, w)
    loss2, grad2 = net.loss_grad_raw(x, t, w)
    assert loss2 == loss and np.array_equal(grad2, grad), "gradients are not deterministic"
    orc = so.Net(flat, C, rounding=True)
    g = {"in": x}
    for l in range(17):
        g[l] = _fetch(net, l)
    dz = {l: _fetch(net, 17 + l) for l in range(14)}
    # head: loss terms and gradients from the GPU's last activation
    h = g[13]
    out = h @ orc.wh.T + orc.bh
    wb = (w / B)[..., None]
    e = out - t
    assert np.isclose(loss, np.sum(wb * e * e), rtol=1e-4)
    gout = 2 * wb * e
    layers, _, _ = so.unflatten(grad, C)
    o = so.num_params(C) - (4 * C + 4)
    np.testing.assert_allclose(grad[o:o + 4 * C].reshape(4, C),
                               np.einsum("bhwj,bhwc->jc", gout, h), rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(grad[o + 4 * C:], gout.sum(axis=(0, 1, 2)), rtol=1e-4, atol=1e-7)
    _assert_bf16_close(dz[13], so.bf16(np.where(h > 0, gout @ orc.wh, 0.0)), "head dh")
    # each convolution: weight and bias gradients, and the data gradient into the layer below
    skip, dpool = {}, {}
    for l in range(13, -1, -1):
        name = so.layer_defs(C)[l][0]
        x1, x2 = _layer_inputs(g, l)
        xin = x1 if x2 is None else np.concatenate([x1, so.up2(x2)], axis=-1)
        dw = so.conv3x3_weight_grad(dz[l], xin)
        db = dz[l].sum(axis=(0, 1, 2))
        gw, gb = layers[l]
        assert np.linalg.norm(gw - dw) <= 1e-4 * np.linalg.norm(dw) + 1e-12, name
        assert np.linalg.norm(gb - db) <= 1e-4 * np.linalg.norm(db) + 1e-12, name
        if l == 0:
            break
        dx = so.conv3x3_input_grad(dz[l], orc.wconv[l])
        if l in (13, 11, 9, 7, 5, 3, 1):
            _assert_bf16_close(dz[l - 1], so.bf16(np.where(g[l - 1] > 0, dx, 0.0)), f"dgrad {name}")
        elif l in (12, 10, 8):
            c1 = so.layer_defs(C)[l][1]
            coarse = {12: 11, 10: 9, 8: 7}[l]
            skip[l] = so.bf16(dx[..., :c1])
            up = so.bf16(dx[..., c1:])
            want = so.bf16(np.where(g[coarse] > 0, so.up2_back(up), 0.0))
            mag = np.where(g[coarse] > 0, so.up2_back(np.abs(up)), 0.0)
            _assert_bf16_close(dz[coarse], want, f"dgrad+upsample {name}", steps=2, frac=0.03,
                               addends=mag)
        else:
            act = {6: 5, 4: 3, 2: 1}[l]
            s = skip[{6: 8, 4: 10, 2: 12}[l]]
            want = so.bf16(np.where(g[act] > 0, s + 0.25 * so.up2(so.bf16(dx)), 0.0))
            mag = np.where(g[act] > 0, np.abs(s) + 0.25 * so.up2(np.abs(so.bf16(dx))), 0.0)
            _assert_bf16_close(dz[act], want, f"dgrad+pool {name}", steps=2, frac=0.03,
                               addends=mag)
    # end to end against the rounded oracle network
    lo, go = orc.loss_grad(x, t, w)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    for l, ((gw, gb), (ow, ob)) in enumerate(zip(layers, so.unflatten(go, C)[0])):
        rel = np.linalg.norm(gw - ow) / np.linalg.norm(ow)
        assert rel < 5e-2, (so.layer_defs(C)[l][0], rel)


def _analytic_fields(p, Hh, Ww):
    """A smooth field family with a bow-shock-like bump and a wake, as a function of the case
    parameters, for the training test (physical units)."""
    nf, uf, tf = so.freestream(p)
    xs = np.linspace(-1, 1, Hh)[:, None]
    ys = np.linspace(-1, 1, Ww)[None, :]
    a = np.radians(p.attack_angle_deg)
    xr = np.cos(a) * xs - np.sin(a) * ys
    yr = np.sin(a) * xs + np.cos(a) * ys
    bump = np.exp(-((xr + 0.15) ** 2 + yr ** 2) / 0.02)
    wake = np.exp(-(yr ** 2) / 0.01) * (xr > 0.1)
    strength = 0.4 * p.mach / 4.0 * (1.0 + 0.5 * (p.wall_model == 0))
    n = nf * (1.0 + strength * bump) * (1.0 - 0.6 * wake)
    u = uf * (1.0 - 0.7 * bump - 0.3 * wake)
    t = tf * (1.0 + 2.0 * strength * bump)
    f = np.stack([n, u * np.cos(a), -u * np.sin(a), t]).astype(np.float32)
    f[:, so.raster_mask(p, Hh, Ww)] = 0.0
    return f


def test_training_fits_an_analytic_family_and_is_reproducible():
    from dsmc_erlend_amd import default_params
    from dsmc_erlend_amd.surrogate import Surrogate
    rng = np.random.default_rng(11)
    Hh = Ww = 32

    def case(i):
        return default_params(attack_angle_deg=float(rng.uniform(-30, 30)),
                              mach=float(rng.uniform(2, 8)), wall_model=i % 3,
                              n_free=float(10 ** rng.uniform(16, 18)))

    train = [case(i) for i in range(48)]
    test = [case(i) for i in range(8)]
    ftrain = np.stack([_analytic_fields(p, Hh, Ww) for p in train])
    ftest = np.stack([_analytic_fields(p, Hh, Ww) for p in test])
    weights = []
    for rep in range(2):
        with Surrogate(height=Hh, width=Ww, max_batch=8, seed=5) as net:
            net.set_dataset(train, ftrain)
            before = net.evaluate(test, ftest)
            hist = net.train(epochs=40, batch=8, lr=2e-3, seed=3)
            after = net.evaluate(test, ftest)
            weights.append(net.weights())
            pred = net.predict(test)
    assert hist.size == 40 * 6 and np.all(np.isfinite(hist))
    assert hist[-12:].mean() < 0.1 * hist[:3].mean(), (hist[:3], hist[-12:])
    assert after < 0.2 * before, (before, after)
    assert np.array_equal(weights[0], weights[1]), "training is not reproducible"
    # predictions come back in physical units, zero density inside the obstacle
    mask = so.raster_mask(test[0], Hh, Ww)
    assert np.all(pred[0, 0][mask] == 0) and np.all(pred[:, 0][:, ~mask] > 0)
    nf, _, _ = so.freestream(test[0])
    assert 0.2 * nf < np.median(pred[0, 0][~mask]) < 5 * nf
