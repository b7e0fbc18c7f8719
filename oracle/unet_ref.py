"""NumPy oracle of the learned coarse-level correction network (TEST INFRASTRUCTURE ONLY:
imported by tests/ as the checker of the HIP kernels, never by the product path).

The network has no counterpart in the reference; this file restates its definition so the
GPU forward / backward kernels can be checked against it:

* conv3d(x, W, b, stride): y[o] = b + sum_t W_t^T x[stride*o + d_t], taps d_t in {-1,0,1}^3
  with t = (dx+1) + 3(dy+1) + 9(dz+1), zero padding, output box ceil(n / stride).
* upsample(x, fine): node-aligned trilinear interpolation (fine 2I = coarse I, odd fine
  nodes halfway, the missing far neighbour at an even-sized fine box counts as 0).
* U-Net 16/32/64 with skip concatenation [upsampled | encoder], ReLU after every
  convolution except the last, FP32 (3-channel) output.
* loss = 1/B sum_b |m (y - t)|^2 / |m t|^2.

Tensors are [B, Z, Y, X, C].  With ``bf16=True`` every value the GPU stores in BF16 is
rounded at the same place (weights, activations, gradients), accumulation is float64.
Pinned by: a brute-force loop convolution, adjoint identities of the convolution and the
upsampling, exactness of the upsampling on linear fields, and central finite differences
of the full loss (tests/test_unet_oracle.py).
"""
import numpy as np

CH = (16, 32, 64)
LAYERS = ("enc0a", "enc0b", "enc1a", "enc1b", "enc2a", "enc2b", "dec1a", "dec1b", "dec0a", "dec0b", "out")


def layer_table(ch=CH):
    c0, c1, c2 = ch
    return [(4, c0, 1, True), (c0, c0, 1, True), (c0, c1, 2, True), (c1, c1, 1, True), (c1, c2, 2, True),
            (c2, c2, 1, True), (c2 + c1, c1, 1, True), (c1, c1, 1, True), (c1 + c0, c0, 1, True),
            (c0, c0, 1, True), (c0, 3, 1, False)]


def n_params(ch=CH):
    return sum(27 * ci * co + co for ci, co, _, _ in layer_table(ch))


def unpack(theta, ch=CH):
    """flat parameter vector -> list of (W [27, cin, cout], b [cout])"""
    out, off = [], 0
    for ci, co, _, _ in layer_table(ch):
        W = theta[off:off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        b = theta[off:off + co]
        off += co
        out.append((W, b))
    return out


def pack(params):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params])


def bf16(a):
    """round float values to the nearest BF16 (ties to even), returned as float64"""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _taps():
    for t in range(27):
        yield t, (t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1)


def out_shape(shape, stride):
    return tuple((n + stride - 1) // stride for n in shape)


def _src_slices(n_in, n_out, d, stride):
    """slices (out, in) such that in = stride*o + d stays inside [0, n_in)"""
    o_lo = 0
    while stride * o_lo + d < 0:
        o_lo += 1
    o_hi = n_out
    while o_hi > o_lo and stride * (o_hi - 1) + d >= n_in:
        o_hi -= 1
    if o_hi <= o_lo:
        return None
    i_lo = stride * o_lo + d
    return slice(o_lo, o_hi), slice(i_lo, i_lo + stride * (o_hi - o_lo - 1) + 1, stride)


def conv3d(x, W, b=None, stride=1):
    B, Z, Y, X, _ = x.shape
    oz, oy, ox = out_shape((Z, Y, X), stride)
    y = np.zeros((B, oz, oy, ox, W.shape[2]))
    for t, (dx, dy, dz) in _taps():
        sz, sy, sx = _src_slices(Z, oz, dz, stride), _src_slices(Y, oy, dy, stride), _src_slices(X, ox, dx, stride)
        if sz is None or sy is None or sx is None:
            continue
        y[:, sz[0], sy[0], sx[0], :] += x[:, sz[1], sy[1], sx[1], :] @ W[t]
    if b is not None:
        y += b
    return y


def conv3d_dgrad(dy, W, in_shape, stride=1):
    """adjoint of conv3d in x"""
    B = dy.shape[0]
    Z, Y, X = in_shape
    oz, oy, ox = dy.shape[1:4]
    dx_ = np.zeros((B, Z, Y, X, W.shape[1]))
    for t, (dx, dyy, dz) in _taps():
        sz, sy, sx = _src_slices(Z, oz, dz, stride), _src_slices(Y, oy, dyy, stride), _src_slices(X, ox, dx, stride)
        if sz is None or sy is None or sx is None:
            continue
        dx_[:, sz[1], sy[1], sx[1], :] += dy[:, sz[0], sy[0], sx[0], :] @ W[t].T
    return dx_


def conv3d_wgrad(x, dy, stride=1):
    B, Z, Y, X, ci = x.shape
    oz, oy, ox, co = dy.shape[1:]
    dW = np.zeros((27, ci, co))
    for t, (dx, dyy, dz) in _taps():
        sz, sy, sx = _src_slices(Z, oz, dz, stride), _src_slices(Y, oy, dyy, stride), _src_slices(X, ox, dx, stride)
        if sz is None or sy is None or sx is None:
            continue
        xs = x[:, sz[1], sy[1], sx[1], :].reshape(-1, ci)
        ds = dy[:, sz[0], sy[0], sx[0], :].reshape(-1, co)
        dW[t] = xs.T @ ds
    return dW


def _up_matrix(nf, nc):
    P = np.zeros((nf, nc))
    for u in range(nf):
        if u % 2 == 0:
            P[u, u // 2] = 1.0
        else:
            P[u, u // 2] = 0.5
            if u // 2 + 1 < nc:
                P[u, u // 2 + 1] = 0.5
    return P


def upsample(x, fine):
    Zf, Yf, Xf = fine
    _, Zc, Yc, Xc, _ = x.shape
    Pz, Py, Px = _up_matrix(Zf, Zc), _up_matrix(Yf, Yc), _up_matrix(Xf, Xc)
    return np.einsum("az,by,cx,nzyxk->nabck", Pz, Py, Px, x, optimize=True)


def upsample_adj(dy, coarse):
    Zc, Yc, Xc = coarse
    _, Zf, Yf, Xf, _ = dy.shape
    Pz, Py, Px = _up_matrix(Zf, Zc), _up_matrix(Yf, Yc), _up_matrix(Xf, Xc)
    return np.einsum("az,by,cx,nabck->nzyxk", Pz, Py, Px, dy, optimize=True)


def forward(params, feat, bf=False, ch=CH):
    """feat [B, Z, Y, X, 4] -> (y [B, Z, Y, X, 3], cache)"""
    r = bf16 if bf else (lambda a: a)
    P = [(r(W), b) for W, b in params]
    c0, c1, c2 = ch
    g0 = feat.shape[1:4]
    g1 = out_shape(g0, 2)
    A = {"feat": r(feat)}

    def conv(li, x, stride, relu=True):
        y = conv3d(x, P[li][0], P[li][1], stride)
        return r(np.maximum(y, 0.0)) if relu else y

    A["e0a"] = conv(0, A["feat"], 1)
    A["e0b"] = conv(1, A["e0a"], 1)
    A["e1a"] = conv(2, A["e0b"], 2)
    A["e1b"] = conv(3, A["e1a"], 1)
    A["e2a"] = conv(4, A["e1b"], 2)
    A["e2b"] = conv(5, A["e2a"], 1)
    A["cat1"] = np.concatenate([r(upsample(A["e2b"], g1)), A["e1b"]], axis=-1)
    A["d1a"] = conv(6, A["cat1"], 1)
    A["d1b"] = conv(7, A["d1a"], 1)
    A["cat0"] = np.concatenate([r(upsample(A["d1b"], g0)), A["e0b"]], axis=-1)
    A["d0a"] = conv(8, A["cat0"], 1)
    A["d0b"] = conv(9, A["d0a"], 1)
    A["y"] = conv(10, A["d0b"], 1, relu=False)
    A["P"] = P
    return A["y"], A


def loss_and_grad(params, feat, tgt, mask, bf=False, ch=CH):
    """returns (loss, list of (dW, db)) with the GPU's rounding points when bf"""
    r = bf16 if bf else (lambda a: a)
    y, A = forward(params, feat, bf, ch)
    P = A["P"]
    B = feat.shape[0]
    m = mask[..., None].astype(np.float64)
    den = np.array([np.sum((m[b] * tgt[b]) ** 2) for b in range(B)])
    sc = np.where(den > 1e-30, 1.0 / np.maximum(den, 1e-300), 0.0) / B
    d = m * (y - tgt)
    loss = float(np.sum(sc[:, None, None, None, None] * d * d))
    gy = r(2.0 * sc[:, None, None, None, None] * d)
    c0, c1, c2 = ch
    g0 = feat.shape[1:4]
    g1 = out_shape(g0, 2)
    g2 = out_shape(g1, 2)
    grads = [None] * 11

    def wb(li, x, dy, stride):
        grads[li] = (conv3d_wgrad(x, dy, stride), dy.sum(axis=(0, 1, 2, 3)))

    def dg(li, dy, shape, stride, mask_from=None, add=None):
        v = conv3d_dgrad(dy, P[li][0], shape, stride)
        if add is not None:
            v = v + add
        if mask_from is not None:
            v = v * (mask_from > 0)
        return r(v)

    g_d0b = dg(10, gy, g0, 1, A["d0b"])
    wb(10, A["d0b"], gy, 1)
    g_d0a = dg(9, g_d0b, g0, 1, A["d0a"])
    wb(9, A["d0a"], g_d0b, 1)
    g_cat0 = dg(8, g_d0a, g0, 1)
    wb(8, A["cat0"], g_d0a, 1)
    g_d1b = r(upsample_adj(g_cat0[..., :c1], g1) * (A["d1b"] > 0))
    g_d1a = dg(7, g_d1b, g1, 1, A["d1a"])
    wb(7, A["d1a"], g_d1b, 1)
    g_cat1 = dg(6, g_d1a, g1, 1)
    wb(6, A["cat1"], g_d1a, 1)
    g_e2b = r(upsample_adj(g_cat1[..., :c2], g2) * (A["e2b"] > 0))
    g_e2a = dg(5, g_e2b, g2, 1, A["e2a"])
    wb(5, A["e2a"], g_e2b, 1)
    g_e1b = dg(4, g_e2a, g1, 2, A["e1b"], g_cat1[..., c2:])
    wb(4, A["e1b"], g_e2a, 2)
    g_e1a = dg(3, g_e1b, g1, 1, A["e1a"])
    wb(3, A["e1a"], g_e1b, 1)
    g_e0b = dg(2, g_e1a, g0, 2, A["e0b"], g_cat0[..., c1:])
    wb(2, A["e0b"], g_e1a, 2)
    g_e0a = dg(1, g_e0b, g0, 1, A["e0a"])
    wb(1, A["e0a"], g_e0b, 1)
    wb(0, A["feat"], g_e0a, 1)
    return loss, grads


def init_params(seed, ch=CH):
    rng = np.random.default_rng(seed)
    out = []
    for ci, co, _, relu in layer_table(ch):
        sd = np.sqrt(2.0 / (27 * ci)) * (1.0 if relu else 0.1)
        out.append((rng.normal(0, sd, (27, ci, co)), np.zeros(co)))
    return out
