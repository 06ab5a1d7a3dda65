"""FP64 numpy restatement of the denoiser network (the checker of the GPU kernels).

TEST INFRASTRUCTURE ONLY.  The network has no reference counterpart; this file
states its mathematics directly: 3x3x3 convolutions with zero padding, taps
ordered t = (dz+1)*9 + (dy+1)*3 + (dx+1), weights [cout][27][cin], ReLU after
every layer but the last, loss = mean squared error.  Two heads:
  residual  prediction = output[:3] + the 3 input channels at resid_ch
  kernel    output[:27] are logits of a 3x3x3 filter per voxel: prediction =
            sum over the taps inside the volume of softmax(logits)_t times the
            3 input channels at resid_ch of the tap's voxel (the softmax runs
            over the in-volume taps only)
Optional BF16 rounding of activations and gradients reproduces the GPU's
storage precision.
"""
import numpy as np

LAYERS = 8
WIDTH = 32


HEADS = ("residual", "kernel")


def shapes(head="residual"):
    last = 16 if head == "residual" else 32
    return [(WIDTH if l < LAYERS - 1 else last, 16 if l == 0 else WIDTH) for l in range(LAYERS)]


def bf16(a):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, head="residual"):
    out, off = [], 0
    for co, ci in shapes(head):
        w = params[off:off + co * 27 * ci].reshape(co, 27, ci).astype(np.float64)
        off += co * 27 * ci
        b = params[off:off + co].astype(np.float64)
        off += co
        out.append((w, b))
    assert off == params.size
    return out


def _shift(xp, tap, d, h, w):
    dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
    return xp[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :]


def conv3d(x, w, b=None, relu=False):
    nb, d, h, wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nb, d, h, wd, w.shape[0]))
    for t in range(27):
        y += _shift(xp, t, d, h, wd) @ w[:, t, :].T
    if b is not None:
        y += b
    if relu:
        y = np.maximum(y, 0)
    return y


def pad_input(x, cin=16):
    nb, d, h, w, c = x.shape
    out = np.zeros((nb, d, h, w, cin))
    out[..., :c] = x
    return out


def _tap_views(rad):
    """The 27 neighbour views of rad [nb, d, h, w, c] and their in-volume masks."""
    nb, d, h, w, _ = rad.shape
    rp = np.pad(rad, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    ones = np.pad(np.ones((nb, d, h, w, 1)), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    views = np.stack([_shift(rp, t, d, h, w) for t in range(27)])     # [27, nb, d, h, w, c]
    valid = np.stack([_shift(ones, t, d, h, w)[..., 0] for t in range(27)])  # [27, nb, d, h, w]
    return views, valid


def kernel_weights(logits, valid):
    """softmax over the in-volume taps; logits [..., 27], valid [27, ...]."""
    z = np.moveaxis(logits[..., :27], -1, 0)
    z = np.where(valid > 0, z, -np.inf)
    z = z - z.max(axis=0, keepdims=True)
    e = np.where(valid > 0, np.exp(z), 0.0)
    return e / e.sum(axis=0, keepdims=True)  # [27, ...]


def forward(params, x, resid_ch, emulate_bf16=False, head="residual"):
    layers = unpack(params, head)
    a = pad_input(np.asarray(x, np.float64))
    if emulate_bf16:
        a = bf16(a)
        layers = [(bf16(w), b) for w, b in layers]
    acts = [a]
    for l, (w, b) in enumerate(layers):
        last = l == LAYERS - 1
        a = conv3d(a, w, b, relu=not last)
        if emulate_bf16 and not last:
            a = bf16(a)
        acts.append(a)
    rad = np.asarray(x, np.float64)[..., resid_ch:resid_ch + 3] if resid_ch >= 0 else None
    if head == "kernel":
        views, valid = _tap_views(rad)
        wts = kernel_weights(a, valid)
        return np.einsum("t...,t...c->...c", wts, views), acts
    return a[..., :3] + (rad if rad is not None else 0.0), acts


def loss_grad(params, x, target, resid_ch, emulate_bf16=False, head="residual"):
    layers = unpack(params, head)
    if emulate_bf16:
        layers = [(bf16(w), b) for w, b in layers]
    pred, acts = forward(params, x, resid_ch, emulate_bf16, head)
    V = pred.shape[0] * pred.shape[1] * pred.shape[2] * pred.shape[3]
    e = pred - np.asarray(target, np.float64)
    loss = float(np.sum(e * e) / (3 * V))
    dy = np.zeros(acts[-1].shape)
    dpred = 2 * e / (3 * V)
    if head == "kernel":
        # pred = sum_t w_t rad_t, w = softmax(z): dL/dz_t = w_t (g_t - sum_s w_s g_s), g_t = dpred . rad_t
        views, valid = _tap_views(np.asarray(x, np.float64)[..., resid_ch:resid_ch + 3])
        wts = kernel_weights(acts[-1], valid)
        g = np.einsum("...c,t...c->t...", dpred, views)
        dz = wts * (g - (wts * g).sum(axis=0, keepdims=True))
        dy[..., :27] = np.moveaxis(dz, 0, -1)
    else:
        dy[..., :3] = dpred
    if emulate_bf16:
        dy = bf16(dy)
    grads = []
    for l in range(LAYERS - 1, -1, -1):
        w, b = layers[l]
        xin = acts[l]
        nb, d, h, wd, ci = xin.shape
        xp = np.pad(xin, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
        dw = np.zeros_like(w)
        dyf = dy.reshape(-1, dy.shape[-1])
        for t in range(27):
            dw[:, t, :] = dyf.T @ _shift(xp, t, d, h, wd).reshape(-1, ci)
        db = dy.sum(axis=(0, 1, 2, 3))
        grads.append((dw, db))
        if l == 0:
            break
        dxp = np.zeros_like(xp)
        for t in range(27):
            dz, dyy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            dxp[:, 1 + dz:1 + dz + d, 1 + dyy:1 + dyy + h, 1 + dx:1 + dx + wd, :] += dy @ w[:, t, :]
        dxin = dxp[:, 1:-1, 1:-1, 1:-1, :]
        dy = dxin * (xin > 0)  # ReLU of layer l-1
        if emulate_bf16:
            dy = bf16(dy)
    grads.reverse()
    flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])
    return loss, flat
