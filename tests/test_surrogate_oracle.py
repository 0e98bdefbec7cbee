"""Pins of the warm-start network's CPU oracle (oracle/surrogate_oracle.py), on the CPU.

The network has no reference counterpart, so the oracle is pinned by mathematics and by an
independent library: bf16 rounding against torch's float -> bfloat16 conversion; the 3x3
convolution, pooling and upsampling against brute-force loops; the whole float64 network's
forward pass and gradient against a separate torch implementation under autograd; and the
gradient against central finite differences. The same numpy code paths, with rounding switched
on, are what the GPU is compared with.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import surrogate_oracle as so  # noqa: E402


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
        np.array([0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -9,
                  -(1.0 + 2.0 ** -8), 3.0e38, 1e-40, -1e-40], dtype=np.float32),
    ]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(so.bf16(x), ref)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert so.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert so.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def _conv_brute(x, w, b):
    B, H, W, ci = x.shape
    co = w.shape[0]
    y = np.zeros((B, H, W, co))
    for n in range(B):
        for i in range(H):
            for j in range(W):
                for tap in range(9):
                    ky, kx = divmod(tap, 3)
                    ii, jj = i + ky - 1, j + kx - 1
                    if 0 <= ii < H and 0 <= jj < W:
                        y[n, i, j] += w[:, tap, :] @ x[n, ii, jj]
    return y + b


def test_conv3x3_matches_brute_force_and_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 7, 3))
    w = rng.normal(size=(4, 9, 3))
    b = rng.normal(size=4)
    np.testing.assert_allclose(so.conv3x3(x, w, b), _conv_brute(x, w, b), rtol=1e-12, atol=1e-12)
    # <conv(x), dz> = <x, conv_input_grad(dz)> = <w, conv_weight_grad(dz, x)>
    dz = rng.normal(size=(2, 5, 7, 4))
    lhs = np.sum(so.conv3x3(x, w) * dz)
    assert np.isclose(lhs, np.sum(x * so.conv3x3_input_grad(dz, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * so.conv3x3_weight_grad(dz, x)), rtol=1e-12)


def test_pool_and_upsample_brute_force():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 6, 4, 3))
    p = so.pool2(x)
    for i in range(3):
        for j in range(2):
            np.testing.assert_allclose(p[:, i, j], x[:, 2 * i:2 * i + 2, 2 * j:2 * j + 2].mean(axis=(1, 2)),
                                       rtol=1e-14)
    u = so.up2(p)
    assert u.shape == x.shape and np.all(u[:, 5, 3] == p[:, 2, 1])
    d = rng.normal(size=x.shape)
    assert np.isclose(np.sum(so.up2(p) * d), np.sum(p * so.up2_back(d)), rtol=1e-13)


def _torch_net_loss(torch, flat, C, cin0, inputs, targets, weights):
    F = torch.nn.functional
    t = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    defs = so.layer_defs(C, cin0)
    params, o = [], 0
    for _, c1, c2, co, _ in defs:
        n = co * 9 * (c1 + c2)
        W = t[o:o + n].reshape(co, 3, 3, c1 + c2).permute(0, 3, 1, 2)
        params.append((W, t[o + n:o + n + co]))
        o += n + co
    wh = t[o:o + 4 * C].reshape(4, C)
    bh = t[o + 4 * C:o + 4 * C + 4]

    def conv(l, x):
        W, b = params[l]
        return F.relu(F.conv2d(x, W, b, padding=1))

    def up(x):
        return F.interpolate(x, scale_factor=2, mode="nearest")

    X = torch.tensor(inputs, dtype=torch.float64).permute(0, 3, 1, 2)
    a = {}
    a[0] = conv(0, X); a[1] = conv(1, a[0])
    a[2] = conv(2, F.avg_pool2d(a[1], 2)); a[3] = conv(3, a[2])
    a[4] = conv(4, F.avg_pool2d(a[3], 2)); a[5] = conv(5, a[4])
    a[6] = conv(6, F.avg_pool2d(a[5], 2)); a[7] = conv(7, a[6])
    a[8] = conv(8, torch.cat([a[5], up(a[7])], 1)); a[9] = conv(9, a[8])
    a[10] = conv(10, torch.cat([a[3], up(a[9])], 1)); a[11] = conv(11, a[10])
    a[12] = conv(12, torch.cat([a[1], up(a[11])], 1)); a[13] = conv(13, a[12])
    out = torch.einsum("bchw,jc->bhwj", a[13], wh) + bh
    w = torch.tensor(weights, dtype=torch.float64)[..., None] / inputs.shape[0]
    e = out - torch.tensor(targets, dtype=torch.float64)
    loss = torch.sum(w * e * e)
    loss.backward()
    return loss.item(), t.grad.numpy(), out.detach().numpy()


def _small_problem(seed, C=4, cin0=5, B=2, H=16, W=8):
    rng = np.random.default_rng(seed)
    flat = np.zeros(so.num_params(C, cin0))
    o = 0
    for _, c1, c2, co, _ in so.layer_defs(C, cin0):
        n = co * 9 * (c1 + c2)
        flat[o:o + n] = rng.normal(size=n) * np.sqrt(2.0 / (9 * (c1 + c2)))
        flat[o + n:o + n + co] = rng.normal(size=co) * 0.1
        o += n + co
    flat[o:] = rng.normal(size=4 * C + 4) * 0.5
    inputs = rng.normal(size=(B, H, W, cin0))
    targets = rng.normal(size=(B, H, W, 4))
    weights = rng.uniform(size=(B, H, W)) * (rng.uniform(size=(B, H, W)) > 0.2)
    return flat, inputs, targets, weights


def test_network_forward_and_gradient_match_torch_autograd():
    torch = pytest.importorskip("torch")
    C, cin0 = 4, 5
    flat, x, t, w = _small_problem(3, C, cin0)
    net = so.Net(flat, C, cin0, rounding=False)
    loss, grad = net.loss_grad(x, t, w)
    tloss, tgrad, tout = _torch_net_loss(torch, flat, C, cin0, x, t, w)
    np.testing.assert_allclose(net.forward(x), tout, rtol=1e-12, atol=1e-12)
    assert np.isclose(loss, tloss, rtol=1e-12)
    np.testing.assert_allclose(grad, tgrad, rtol=1e-9, atol=1e-12 * np.abs(tgrad).max())


def test_network_gradient_matches_finite_differences():
    C, cin0 = 4, 5
    flat, x, t, w = _small_problem(4, C, cin0, B=1, H=8, W=8)
    net = so.Net(flat, C, cin0, rounding=False)
    _, grad = net.loss_grad(x, t, w)
    rng = np.random.default_rng(5)
    # every layer's first weight and bias, plus random entries
    idx, o = [], 0
    for _, c1, c2, co, _ in so.layer_defs(C, cin0):
        n = co * 9 * (c1 + c2)
        idx += [o, o + n]
        o += n + co
    idx += list(range(o, o + 4 * C + 4, 5)) + list(rng.integers(0, flat.size, 30))
    h = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp, _ = so.Net(fp, C, cin0, rounding=False).loss_grad(x, t, w)
        lm, _ = so.Net(fm, C, cin0, rounding=False).loss_grad(x, t, w)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_rounded_network_stays_close_to_exact():
    # bf16 storage perturbs, but does not change, the function: a plausible misplaced rounding
    # (e.g. of a sum) would show as a much larger gap
    C, cin0 = 8, 5
    flat, x, t, w = _small_problem(6, C, cin0)
    exact = so.Net(flat, C, cin0, rounding=False)
    rnd = so.Net(flat, C, cin0, rounding=True)
    le, ge = exact.loss_grad(x, t, w)
    lr, gr = rnd.loss_grad(x, t, w)
    assert abs(lr - le) < 2e-2 * abs(le)
    assert np.linalg.norm(gr - ge) < 5e-2 * np.linalg.norm(ge)


def test_normalize_freestream_field_and_mask():
    from dsmc_erlend_amd import default_params
    p = default_params(nx=64, ny=64, attack_angle_deg=15.0)
    H = W = 64
    nf, uf, tf = so.freestream(p)
    fields = np.zeros((4, H, W))
    fields[0] = nf
    fields[1] = uf * np.cos(np.radians(15.0))
    fields[2] = -uf * np.sin(np.radians(15.0))
    fields[3] = tf
    mask = so.raster_mask(p, H, W)
    # a 4x4 m box on a 40 m domain at 64 cells: centres within (-2, 2) -> 2 * 3 cells per side
    assert mask.sum() == 36
    fields[:, mask] = 0.0
    tg, wt = so.normalize(p, fields)
    ok = ~mask
    assert np.allclose(tg[ok, 0], 0.0) and np.allclose(tg[ok, 3], 0.0)
    assert np.allclose(tg[ok, 1], np.cos(np.radians(15.0)))
    assert np.allclose(tg[ok, 2], -np.sin(np.radians(15.0)))
    assert np.all(wt[mask] == 0) and np.isclose(wt.sum(), 0.25)
