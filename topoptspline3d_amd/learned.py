"""numpy-facing wrappers of the learned coarse-level correction (include/tos_api.h,
"learned coarse-level correction"): the U-Net, its pair dataset, and the calls that
attach it to / log pairs from a design-domain Context."""
import ctypes as C

import numpy as np

from .context import TosError


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class UNet:
    """3-level 3-D conv U-Net (16/32/64 channels) on the GPU."""

    def __init__(self, api, seed=0, channels=(16, 32, 64)):
        self.api = api
        h = C.c_void_p()
        rc = api.unet_create(channels[0], channels[1], channels[2], int(seed), C.byref(h))
        if rc:
            raise TosError(rc, "unet_create", "")
        self.h = h
        n = C.c_int64()
        api.unet_num_params(self.h, C.byref(n))
        self.n_params = n.value

    def _chk(self, rc, where):
        if rc:
            msg = self.api.unet_last_error(self.h)
            raise TosError(rc, where, msg.decode() if msg else "")

    def close(self):
        if self.h:
            self.api.unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self._chk(self.api.unet_get_params(self.h, _p(out, C.c_float)), "unet_get_params")
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        assert theta.size == self.n_params
        self._chk(self.api.unet_set_params(self.h, _p(theta, C.c_float)), "unet_set_params")

    def forward(self, feat):
        """feat [B, Z, Y, X, 4] -> [B, Z, Y, X, 3]"""
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        B, Z, Y, X, _ = feat.shape
        out = np.zeros((B, Z, Y, X, 3), np.float32)
        self._chk(self.api.unet_forward(self.h, B, X, Y, Z, _p(feat, C.c_float), _p(out, C.c_float)),
                  "unet_forward")
        return out

    def loss_grad(self, feat, target, mask):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        B, Z, Y, X, _ = feat.shape
        loss = C.c_double()
        grad = np.zeros(self.n_params, np.float32)
        self._chk(self.api.unet_loss_grad(self.h, B, X, Y, Z, _p(feat, C.c_float), _p(target, C.c_float),
                                          _p(mask, C.c_uint8), C.byref(loss), _p(grad, C.c_float)),
                  "unet_loss_grad")
        return loss.value, grad

    def train(self, ds, first, count, steps, batch=8, lr=1e-3, seed=0):
        lf, ll = C.c_double(), C.c_double()
        self._chk(self.api.unet_train(self.h, ds.h, first, count, steps, batch, lr, int(seed), C.byref(lf),
                                      C.byref(ll)), "unet_train")
        return lf.value, ll.value

    def evaluate(self, ds, first, count):
        v = C.c_double()
        self._chk(self.api.unet_eval(self.h, ds.h, first, count, C.byref(v)), "unet_eval")
        return v.value


class PairDataset:
    """Device-resident (features, target, mask) pairs on one node box."""

    def __init__(self, api, box, capacity):
        self.api = api
        X, Y, Z = box
        h = C.c_void_p()
        rc = api.dataset_create(X, Y, Z, capacity, C.byref(h))
        if rc:
            raise TosError(rc, "dataset_create", "")
        self.h = h
        self.box = (X, Y, Z)
        self.capacity = capacity

    def close(self):
        if self.h:
            self.api.dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = C.c_int()
        self.api.dataset_size(self.h, C.byref(n))
        return n.value

    def get(self, i):
        X, Y, Z = self.box
        f = np.zeros((Z, Y, X, 4), np.float32)
        t = np.zeros((Z, Y, X, 3), np.float32)
        m = np.zeros((Z, Y, X), np.uint8)
        rc = self.api.dataset_get(self.h, i, _p(f, C.c_float), _p(t, C.c_float), _p(m, C.c_uint8))
        if rc:
            raise TosError(rc, "dataset_get", "")
        return f, t, m


def unet_flops(box, channels=(16, 32, 64)):
    """Algorithmic FLOPs of one forward pass on a node box (X, Y, Z): 2*27*cin*cout per
    output voxel of every layer (enc0a/b, enc1a/b, enc2a/b at full, half and quarter
    resolution, the two decoder blocks on the skip concatenations, and the 3-channel
    output layer); half a box is ceil(n/2) per axis.  A training step (forward, data and
    weight gradients) is about three times this."""
    c0, c1, c2 = channels
    X, Y, Z = box
    h = lambda g: tuple((n + 1) // 2 for n in g)
    V0 = X * Y * Z
    g1 = h(box)
    V1 = g1[0] * g1[1] * g1[2]
    g2 = h(g1)
    V2 = g2[0] * g2[1] * g2[2]
    conv = lambda ci, co, V: 2 * 27 * ci * co * V
    return (conv(4, c0, V0) + conv(c0, c0, V0) + conv(c0, c1, V1) + conv(c1, c1, V1) + conv(c1, c2, V2)
            + conv(c2, c2, V2) + conv(c2 + c1, c1, V1) + conv(c1, c1, V1) + conv(c1 + c0, c0, V0)
            + conv(c0, c0, V0) + conv(c0, 3, V0))


def level_box(ctx, level):
    """node box (X, Y, Z) of multigrid level `level` of a Context"""
    ex, ey, ez = ctx.info()["level_dims"][level]
    return (ex + 1, ey + 1, ez + 1)


def attach(ctx, net, level=2):
    ctx._chk(ctx.api.ctx_learn_attach(ctx.h, net.h if net is not None else None, level,
                                       1 if net is not None else 0), "ctx_learn_attach")


def log_pairs(ctx, ds, level=2, n_inner=8, stride=1):
    ctx._chk(ctx.api.ctx_learn_log(ctx.h, ds.h if ds is not None else None, level, n_inner, stride), "ctx_learn_log")


def stats(ctx):
    """(corrections tried, kept, mean energy-optimal step of those kept)"""
    t, a, r = C.c_int(), C.c_int(), C.c_double()
    ctx._chk(ctx.api.ctx_learn_stats(ctx.h, C.byref(t), C.byref(a), C.byref(r)), "ctx_learn_stats")
    return t.value, a.value, r.value
