"""The learned warm start: a ctypes view of include/dsmc_surrogate.h.

`Surrogate` owns one network on one GPU (weights, activations and the training set in device
memory). Its kernels are the library's own bf16 MFMA convolutions; nothing here computes on the
host except argument marshalling. `normalize` is the host-side target normalisation the library
uses for training, exposed for tests.
"""
import ctypes as C

import numpy as np

from . import DsmcError, load_library
from ._abi import Params

IN_CHANNELS = 32
OUT_CHANNELS = 4
NUM_FEATURES = 9

# (name, first-input channels, upsampled-input channels, output channels, level) for base C;
# level l works on an (H >> l) x (W >> l) raster.
LAYER_NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b",
               "d2a", "d2b", "d1a", "d1b", "d0a", "d0b"]


def layer_defs(C_, cin0=IN_CHANNELS):
    c = C_
    return [
        ("e0a", cin0, 0, c, 0), ("e0b", c, 0, c, 0),
        ("e1a", c, 0, 2 * c, 1), ("e1b", 2 * c, 0, 2 * c, 1),
        ("e2a", 2 * c, 0, 4 * c, 2), ("e2b", 4 * c, 0, 4 * c, 2),
        ("e3a", 4 * c, 0, 8 * c, 3), ("e3b", 8 * c, 0, 8 * c, 3),
        ("d2a", 4 * c, 8 * c, 4 * c, 2), ("d2b", 4 * c, 0, 4 * c, 2),
        ("d1a", 2 * c, 4 * c, 2 * c, 1), ("d1b", 2 * c, 0, 2 * c, 1),
        ("d0a", c, 2 * c, c, 0), ("d0b", c, 0, c, 0),
    ]


class NetConfig(C.Structure):
    _fields_ = [("height", C.c_int32), ("width", C.c_int32), ("base_channels", C.c_int32),
                ("max_batch", C.c_int32), ("seed", C.c_uint64)]


_fp = C.POINTER(C.c_float)
_declared = set()


def _lib():
    lib = load_library()
    if id(lib) in _declared:
        return lib
    vp = C.c_void_p
    pp = C.POINTER(Params)
    sig = {
        "dsmc_net_config_default": (None, [C.POINTER(NetConfig)]),
        "dsmc_net_create": (C.c_int, [C.POINTER(NetConfig), C.c_int, C.POINTER(C.c_void_p)]),
        "dsmc_net_destroy": (None, [vp]),
        "dsmc_net_last_error": (C.c_char_p, [vp]),
        "dsmc_net_num_params": (C.c_int64, [vp]),
        "dsmc_net_get_weights": (C.c_int, [vp, _fp, C.c_size_t]),
        "dsmc_net_set_weights": (C.c_int, [vp, _fp, C.c_size_t]),
        "dsmc_net_encode": (C.c_int, [vp, pp, C.c_int, _fp]),
        "dsmc_net_normalize": (C.c_int, [C.POINTER(NetConfig), pp, C.c_int, _fp, _fp, _fp]),
        "dsmc_net_forward_raw": (C.c_int, [vp, _fp, C.c_int, _fp]),
        "dsmc_net_get_activation": (C.c_int, [vp, C.c_int, _fp, C.c_size_t,
                                               C.POINTER(C.c_size_t)]),
        "dsmc_net_loss_grad_raw": (C.c_int, [vp, _fp, _fp, _fp, C.c_int, _fp, _fp, C.c_size_t]),
        "dsmc_net_set_dataset": (C.c_int, [vp, pp, _fp, C.c_int]),
        "dsmc_net_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _fp,
                                     C.c_size_t]),
        "dsmc_net_evaluate": (C.c_int, [vp, pp, _fp, C.c_int, _fp]),
        "dsmc_net_predict": (C.c_int, [vp, pp, C.c_int, _fp]),
        "dsmc_net_last_train_stats": (C.c_int, [vp, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _declared.add(id(lib))
    return lib


def default_config(**kw):
    cfg = NetConfig()
    _lib().dsmc_net_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(_fp)


def _params_array(cases):
    cases = list(cases)
    return (Params * len(cases))(*cases), len(cases)


def normalize(cfg, cases, fields):
    """Training targets [n, H, W, 4] and loss weights [n, H, W] of physical fields
    [n, 4, H, W] (host only)."""
    arr, n = _params_array(cases)
    f = _f32(fields)
    t = np.zeros((n, cfg.height, cfg.width, 4), np.float32)
    w = np.zeros((n, cfg.height, cfg.width), np.float32)
    rc = _lib().dsmc_net_normalize(C.byref(cfg), arr, n, _ptr(f), _ptr(t), _ptr(w))
    if rc:
        raise DsmcError(rc, "normalize failed")
    return t, w


class Surrogate:
    """The warm-start network on one GPU."""

    def __init__(self, cfg=None, device=0, **kw):
        self.lib = _lib()
        self.cfg = cfg if cfg is not None else default_config(**kw)
        h = C.c_void_p()
        rc = self.lib.dsmc_net_create(C.byref(self.cfg), device, C.byref(h))
        self.h = h
        if rc:
            msg = self.lib.dsmc_net_last_error(h).decode() if h.value else "create failed"
            self.close()
            raise DsmcError(rc, msg)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.dsmc_net_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise DsmcError(rc, self.lib.dsmc_net_last_error(self.h).decode())

    @property
    def shape(self):
        return self.cfg.height, self.cfg.width

    def num_params(self):
        return int(self.lib.dsmc_net_num_params(self.h))

    def weights(self):
        out = np.zeros(self.num_params(), np.float32)
        self._check(self.lib.dsmc_net_get_weights(self.h, _ptr(out), out.size))
        return out

    def set_weights(self, w):
        w = _f32(w)
        self._check(self.lib.dsmc_net_set_weights(self.h, _ptr(w), w.size))

    def encode(self, cases):
        arr, n = _params_array(cases)
        out = np.zeros((n, self.cfg.height, self.cfg.width, IN_CHANNELS), np.float32)
        self._check(self.lib.dsmc_net_encode(self.h, arr, n, _ptr(out)))
        return out

    def forward_raw(self, inputs):
        x = _f32(inputs)
        out = np.zeros(x.shape[:3] + (OUT_CHANNELS,), np.float32)
        self._check(self.lib.dsmc_net_forward_raw(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def activation(self, layer):
        """bf16 activation of layer (0..13, 14..16 the pooled tensors) of the last forward;
        17..30 the output gradients of convolutions 0..13 of the last loss_grad_raw."""
        cnt = C.c_size_t(0)
        self._check(self.lib.dsmc_net_get_activation(self.h, layer, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, np.float32)
        self._check(self.lib.dsmc_net_get_activation(self.h, layer, _ptr(out), out.size,
                                                     C.byref(cnt)))
        return out

    def loss_grad_raw(self, inputs, targets, weights):
        x, t, w = _f32(inputs), _f32(targets), _f32(weights)
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float(0.0)
        self._check(self.lib.dsmc_net_loss_grad_raw(self.h, _ptr(x), _ptr(t), _ptr(w),
                                                    x.shape[0], C.byref(loss), _ptr(g), g.size))
        return float(loss.value), g

    def set_dataset(self, cases, fields):
        arr, n = _params_array(cases)
        f = _f32(fields)
        self._check(self.lib.dsmc_net_set_dataset(self.h, arr, _ptr(f), n))
        self._ndata = n

    def train(self, epochs, batch=8, lr=1e-3, seed=0):
        """Returns the per-step loss history."""
        steps = epochs * (self._ndata // batch)
        hist = np.zeros(max(steps, 1), np.float32)
        self._check(self.lib.dsmc_net_train(self.h, int(epochs), int(batch), float(lr),
                                            int(seed), _ptr(hist), hist.size))
        return hist[:steps]

    def evaluate(self, cases, fields):
        arr, n = _params_array(cases)
        f = _f32(fields)
        loss = C.c_float(0.0)
        self._check(self.lib.dsmc_net_evaluate(self.h, arr, _ptr(f), n, C.byref(loss)))
        return float(loss.value)

    def predict(self, cases):
        """Physical fields [n, 4, H, W] (n, ux, uy, T)."""
        arr, n = _params_array(cases)
        out = np.zeros((n, 4, self.cfg.height, self.cfg.width), np.float32)
        self._check(self.lib.dsmc_net_predict(self.h, arr, n, _ptr(out)))
        return out

    def last_train_stats(self):
        ms, fl = C.c_double(0), C.c_double(0)
        self._check(self.lib.dsmc_net_last_train_stats(self.h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value
