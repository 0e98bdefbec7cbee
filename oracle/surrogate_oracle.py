"""CPU oracle of the warm-start network (test infrastructure only; nothing in the product
imports it).

The network has no counterpart in the reference, which always starts from the freestream fill
(src/Initialization.cpp:290-313). This module restates it in numpy float64 so that the GPU's
bf16 MFMA kernels can be checked:

- `rounding=True` rounds to bf16 (round to nearest even) at exactly the points where the GPU
  stores bf16: inputs, convolution weights, every activation, the pooled tensors, and every
  activation gradient. Sums (convolutions, weight and bias gradients, the head) are exact
  float64 here and FP32 on the GPU, so the two agree to accumulation rounding, plus the rare
  element whose bf16 rounding that difference flips.
- `rounding=False` is the smooth float64 network; the tests pin its backward pass by central
  finite differences and by torch.autograd, and its forward pass by brute-force loops and by
  torch's conv2d.

Layouts follow the library: NHWC tensors [B][H][W][C] with H along x, weights [cout][9][cin]
with tap = 3*ky + kx (ky along H), the flat parameter vector per layer W then b, then the head
Wh [4][C], bh [4].
"""
import numpy as np


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    x = np.ascontiguousarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_defs(C, cin0=32):
    """(name, c1, c2, cout, level) of the 14 convolutions (c2: channels read through the 2x
    nearest upsample of the decoder's concat)."""
    c = C
    return [
        ("e0a", cin0, 0, c, 0), ("e0b", c, 0, c, 0),
        ("e1a", c, 0, 2 * c, 1), ("e1b", 2 * c, 0, 2 * c, 1),
        ("e2a", 2 * c, 0, 4 * c, 2), ("e2b", 4 * c, 0, 4 * c, 2),
        ("e3a", 4 * c, 0, 8 * c, 3), ("e3b", 8 * c, 0, 8 * c, 3),
        ("d2a", 4 * c, 8 * c, 4 * c, 2), ("d2b", 4 * c, 0, 4 * c, 2),
        ("d1a", 2 * c, 4 * c, 2 * c, 1), ("d1b", 2 * c, 0, 2 * c, 1),
        ("d0a", c, 2 * c, c, 0), ("d0b", c, 0, c, 0),
    ]


def num_params(C, cin0=32):
    n = sum(co * 9 * (c1 + c2) + co for _, c1, c2, co, _ in layer_defs(C, cin0))
    return n + 4 * C + 4


def unflatten(flat, C, cin0=32):
    """Flat parameters -> ([(W [cout,9,cin], b [cout])] * 14, Wh [4,C], bh [4]) as float64."""
    flat = np.asarray(flat, dtype=np.float64)
    layers, o = [], 0
    for _, c1, c2, co, _ in layer_defs(C, cin0):
        n = co * 9 * (c1 + c2)
        layers.append((flat[o:o + n].reshape(co, 9, c1 + c2), flat[o + n:o + n + co]))
        o += n + co
    wh = flat[o:o + 4 * C].reshape(4, C)
    bh = flat[o + 4 * C:o + 4 * C + 4]
    assert o + 4 * C + 4 == flat.size
    return layers, wh, bh


def flatten(grads, gwh, gbh):
    parts = []
    for gw, gb in grads:
        parts += [gw.ravel(), gb.ravel()]
    parts += [gwh.ravel(), gbh.ravel()]
    return np.concatenate(parts)


# ---------------------------------------------------------------------------------------
# The operators. x: [B, H, W, C]; taps (ky, kx) in 0..2 with offset (ky - 1, kx - 1).
def _pad(x):
    B, H, W, C = x.shape
    p = np.zeros((B, H + 2, W + 2, C))
    p[:, 1:-1, 1:-1] = x
    return p


def conv3x3(x, w, b=None):
    """y[p, co] = sum_{tap, ci} x[p + off(tap), ci] w[co, tap, ci] (+ b[co]); zero padding."""
    B, H, W, _ = x.shape
    xp = _pad(x)
    y = np.zeros((B, H, W, w.shape[0]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        y += xp[:, ky:ky + H, kx:kx + W, :] @ w[:, tap, :].T
    if b is not None:
        y += b
    return y


def conv3x3_input_grad(dz, w):
    """Gradient of conv3x3 with respect to x: dx[p, ci] = sum_tap dz[p - off(tap)] w[:, tap, ci]."""
    B, H, W, _ = dz.shape
    dx = np.zeros((B, H + 2, W + 2, w.shape[2]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dx[:, ky:ky + H, kx:kx + W, :] += dz @ w[:, tap, :]
    return dx[:, 1:-1, 1:-1, :]


def conv3x3_weight_grad(dz, x):
    """dw[co, tap, ci] = sum_p dz[p, co] x[p + off(tap), ci]."""
    B, H, W, ci = x.shape
    co = dz.shape[-1]
    xp = _pad(x)
    dzf = dz.reshape(-1, co)
    dw = np.zeros((co, 9, ci))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dw[:, tap, :] = dzf.T @ xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, ci)
    return dw


def pool2(x):
    """2x2 average: ((x00 + x01) + x10) + x11, times 1/4."""
    return 0.25 * (((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2])


def up2(x):
    """Nearest-neighbour 2x upsample."""
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_back(d):
    """Adjoint of up2: the sum of each 2x2 block."""
    return ((d[:, 0::2, 0::2] + d[:, 0::2, 1::2]) + d[:, 1::2, 0::2]) + d[:, 1::2, 1::2]


# ---------------------------------------------------------------------------------------
class Net:
    """The encoder-decoder on float64 with optional bf16 rounding at the GPU's storage points."""

    def __init__(self, flat, C, cin0=32, rounding=True):
        self.C, self.cin0, self.rounding = C, cin0, rounding
        self.layers, self.wh, self.bh = unflatten(flat, C, cin0)
        r = self._r
        self.wconv = [r(w) for w, _ in self.layers]
        self.defs = layer_defs(C, cin0)

    def _r(self, a):
        return bf16(a) if self.rounding else np.asarray(a, dtype=np.float64)

    # --- single layers (also used to check the GPU layer by layer) ---
    def conv_layer(self, l, x1, x2=None):
        x = x1 if x2 is None else np.concatenate([x1, up2(x2)], axis=-1)
        z = conv3x3(x, self.wconv[l], self.layers[l][1])
        return self._r(np.maximum(z, 0.0))

    def pool(self, x):
        return self._r(pool2(x))

    def forward(self, inputs):
        r = self._r
        a = {}
        x = r(inputs)
        a["in"] = x
        a[0] = self.conv_layer(0, x)
        a[1] = self.conv_layer(1, a[0])
        a["p0"] = self.pool(a[1])
        a[2] = self.conv_layer(2, a["p0"])
        a[3] = self.conv_layer(3, a[2])
        a["p1"] = self.pool(a[3])
        a[4] = self.conv_layer(4, a["p1"])
        a[5] = self.conv_layer(5, a[4])
        a["p2"] = self.pool(a[5])
        a[6] = self.conv_layer(6, a["p2"])
        a[7] = self.conv_layer(7, a[6])
        a[8] = self.conv_layer(8, a[5], a[7])
        a[9] = self.conv_layer(9, a[8])
        a[10] = self.conv_layer(10, a[3], a[9])
        a[11] = self.conv_layer(11, a[10])
        a[12] = self.conv_layer(12, a[1], a[11])
        a[13] = self.conv_layer(13, a[12])
        out = a[13] @ self.wh.T + self.bh
        self.a = a
        return out

    def layer_inputs(self, l):
        a = self.a
        src = {0: ("in", None), 2: ("p0", None), 4: ("p1", None), 6: ("p2", None),
               8: (5, 7), 10: (3, 9), 12: (1, 11)}
        k1, k2 = src.get(l, (l - 1, None))
        return a[k1], (a[k2] if k2 is not None else None)

    def loss_grad(self, inputs, targets, weights):
        """Mean over the batch of sum_p w_p sum_j (out - t)^2, and the flat gradient."""
        r = self._r
        B = inputs.shape[0]
        out = self.forward(inputs)
        a = self.a
        w = np.asarray(weights, dtype=np.float64)[..., None] / B
        e = out - targets
        loss = float(np.sum(w * e * e))
        g = 2.0 * w * e                                   # [B, H, W, 4]
        h = a[13]
        gwh = np.einsum("bhwj,bhwc->jc", g, h)
        gbh = g.sum(axis=(0, 1, 2))
        dz = {13: r(np.where(h > 0, g @ self.wh, 0.0))}
        grads = [None] * 14
        C = self.C

        def mask(v, act):
            return np.where(act > 0, v, 0.0)

        for l in range(13, -1, -1):
            x1, x2 = self.layer_inputs(l)
            x = x1 if x2 is None else np.concatenate([x1, up2(x2)], axis=-1)
            grads[l] = (conv3x3_weight_grad(dz[l], x), dz[l].sum(axis=(0, 1, 2)))
            if l == 0:
                break
            dx = conv3x3_input_grad(dz[l], self.wconv[l])
            if l in (13, 11, 9, 7, 5, 3, 1):
                dz[l - 1] = r(mask(dx, a[l - 1]))
            elif l in (12, 10, 8):
                c1 = self.defs[l][1]
                skip, up = r(dx[..., :c1]), r(dx[..., c1:])
                coarse = {12: 11, 10: 9, 8: 7}[l]
                dz[coarse] = r(mask(up2_back(up), a[coarse]))
                dz[("skip", l)] = skip
            else:  # 6, 4, 2: the input is a pooled tensor
                dpool = r(dx)
                level_act = {6: 5, 4: 3, 2: 1}[l]
                dskip = dz[("skip", {6: 8, 4: 10, 2: 12}[l])]
                dz[level_act] = r(mask(dskip + 0.25 * up2(dpool), a[level_act]))
        self.dz = {k: v for k, v in dz.items() if isinstance(k, int)}
        return loss, flatten(grads, gwh, gbh)


# ---------------------------------------------------------------------------------------
# Inputs and targets, as the library builds them from a case's parameters.
KB = 1.380662E-23   # the reference's Boltzmann constant (include/Constants.h)
NA = 6.022045E+23   # and Avogadro's number


def freestream(p):
    m = p.molar_mass / NA
    return p.n_free, p.mach * np.sqrt(5.0 / 3.0 * KB * p.t_free / m), p.t_free


def features(p):
    an, at = 1.0, 1.0
    if p.wall_model == 1:
        an = at = 0.0
    elif p.wall_model == 2:
        an, at = p.acc_n, p.acc_t
    return np.array([(p.mach - 5.0) / 3.0, p.attack_angle_deg / 30.0, np.log10(p.n_free) - 17.0,
                     p.wall_model == 0, p.wall_model == 1, p.wall_model == 2, an, at,
                     p.t_wall / p.t_free - 1.5], dtype=np.float32)


def raster_mask(p, H, W):
    """Cell-centre-in-obstacle raster in float32 arithmetic: [H, W] bool."""
    f32 = np.float32
    fx = (np.arange(H, dtype=f32) + f32(0.5)) / f32(H)
    fy = (np.arange(W, dtype=f32) + f32(0.5)) / f32(W)
    X = f32(p.x1) + (f32(p.x2) - f32(p.x1)) * fx
    Y = f32(p.y1) + (f32(p.y2) - f32(p.y1)) * fy
    bx0, bx1 = f32(p.box_x - 0.5 * p.box_width), f32(p.box_x + 0.5 * p.box_width)
    by0, by1 = f32(p.box_y - 0.5 * p.box_height), f32(p.box_y + 0.5 * p.box_height)
    return ((X > bx0) & (X < bx1))[:, None] & ((Y > by0) & (Y < by1))[None, :]


def encode(p, H, W, cin0=32):
    x = np.zeros((H, W, cin0), np.float32)
    x[..., 0] = raster_mask(p, H, W)
    f32 = np.float32
    x[..., 1] = (f32(2) * ((np.arange(H, dtype=f32) + f32(0.5)) / f32(H)) - f32(1))[:, None]
    x[..., 2] = (f32(2) * ((np.arange(W, dtype=f32) + f32(0.5)) / f32(W)) - f32(1))[None, :]
    x[..., 3:12] = features(p)
    return bf16(x)


def normalize(p, fields):
    """Targets [H, W, 4] and weights [H, W] of physical fields [4, H, W]."""
    n, ux, uy, t = (np.asarray(f, dtype=np.float64) for f in fields)
    H, W = n.shape
    nf, uf, tf = freestream(p)
    ok = ~raster_mask(p, H, W) & (n > 0) & (t > 0)
    tg = np.zeros((H, W, 4))
    tg[..., 0] = np.log(np.maximum(n, 1e-3 * nf) / nf)
    tg[..., 1] = ux / uf
    tg[..., 2] = uy / uf
    tg[..., 3] = np.log(np.maximum(t, 1e-3 * tf) / tf)
    tg[~ok] = 0.0
    w = ok / (4.0 * max(int(ok.sum()), 1))
    return tg, w
