"""CPU oracle of the frozen-phonon surrogate network (numpy, float64).

TEST INFRASTRUCTURE ONLY: the checker that the HIP surrogate kernels
(multem_amd/csrc/surrogate.hip) are compared against.  Only tests/ and the
benchmark's CPU leg use it; the library never calls it.

The surrogate has no counterpart in the reference: it maps a few-configuration
frozen-phonon ADF image plus the projected-potential map of the specimen to
the many-configuration ADF image.  Network (all convolutions 3x3 with
periodic padding, since a periodic scan gives a periodic image):

    h0 = relu(conv(x, W0) + b0)                          x: [H, W, 2]
    for k in 1..n_blocks:
        u  = relu(conv(h, Wa_k) + ba_k)
        h  = h + conv(u, Wb_k) + bb_k
    y  = x[..., 0] + conv(h, Wh) + bh                    y: [H, W]
    loss = mean over batch and pixels of (y - t)^2

Images are NHWC; a weight tensor is [cout][3][3][cin] (ky, kx, cin is the
reduction order of the GPU kernels).  conv(x, W)[p, co] =
sum_{ky,kx,ci} x[(py+ky-1) mod H, (px+kx-1) mod W, ci] * W[co, ky, kx, ci].

rounding="bf16" emulates the GPU's arithmetic: every convolution operand
(activation, weight, incoming gradient) is rounded to bfloat16 (round to
nearest even) before the product; products are summed exactly here (FP32 on
the GPU).  rounding="none" is the exact FP64 network, used for the
finite-difference gradient checks that pin the backward pass.
"""
import numpy as np


def bf16(a):
    """Round float values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(f)
    if nan.any():
        out = np.where(nan, f, out)
    return out.astype(np.float64)


def _rnd(a, rounding):
    return bf16(a) if rounding == "bf16" else np.asarray(a, dtype=np.float64)


def layer_shapes(n_blocks=6, ch=32, cin=2):
    """(name, weight shape [cout,3,3,cin], bias shape) in parameter order."""
    s = [("W0", (ch, 3, 3, cin), (ch,))]
    for k in range(n_blocks):
        s.append(("Wa%d" % k, (ch, 3, 3, ch), (ch,)))
        s.append(("Wb%d" % k, (ch, 3, 3, ch), (ch,)))
    s.append(("Wh", (1, 3, 3, ch), (1,)))
    return s


def n_params(n_blocks=6, ch=32, cin=2):
    return sum(int(np.prod(w)) + int(np.prod(b)) for _, w, b in layer_shapes(n_blocks, ch, cin))


def unpack(theta, n_blocks=6, ch=32, cin=2):
    """Flat parameter vector -> list of (W, b).  Order: per layer W then b."""
    out, o = [], 0
    for _, ws, bs in layer_shapes(n_blocks, ch, cin):
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out.append((theta[o:o + nw].reshape(ws), theta[o + nw:o + nw + nb]))
        o += nw + nb
    return out


def init_params(seed, n_blocks=6, ch=32, cin=2):
    """He-uniform weights; the second conv of each block and the head start
    small so that the initial network is close to the identity on channel 0."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, ws, bs in layer_shapes(n_blocks, ch, cin):
        fan_in = ws[1] * ws[2] * ws[3]
        lim = np.sqrt(6.0 / fan_in)
        if name.startswith("Wb") or name == "Wh":
            lim *= 0.1
        parts.append(rng.uniform(-lim, lim, size=int(np.prod(ws))))
        parts.append(np.zeros(int(np.prod(bs))))
    return np.concatenate(parts).astype(np.float32)


def conv3x3(x, W, b=None, rounding="none"):
    """x: [B,H,W,cin], W: [cout,3,3,cin] -> [B,H,W,cout] (periodic padding)."""
    xr = _rnd(x, rounding)
    Wr = _rnd(W, rounding)
    B, H, Wd, cin = xr.shape
    out = np.zeros((B, H, Wd, W.shape[0]))
    for ky in range(3):
        for kx in range(3):
            xs = np.roll(xr, shift=(1 - ky, 1 - kx), axis=(1, 2))   # xs[p] = x[p + k - 1]
            out += xs @ Wr[:, ky, kx, :].T
    if b is not None:
        out += np.asarray(b, dtype=np.float64)
    return out


def conv3x3_grad_input(dy, W, rounding="none"):
    """Gradient of conv3x3 w.r.t. its input: dx[q] = sum_k dy[q - k + 1] W[k]^T."""
    dyr = _rnd(dy, rounding)
    Wr = _rnd(W, rounding)
    dx = np.zeros(dy.shape[:3] + (W.shape[3],))
    for ky in range(3):
        for kx in range(3):
            ds = np.roll(dyr, shift=(ky - 1, kx - 1), axis=(1, 2))  # ds[q] = dy[q - k + 1]
            dx += ds @ Wr[:, ky, kx, :]
    return dx


def conv3x3_grad_weight(x, dy, rounding="none"):
    """dW[co, ky, kx, ci] = sum_p dy[p, co] x[p + k - 1, ci]; db[co] = sum_p dy[p, co]."""
    xr = _rnd(x, rounding)
    dyr = _rnd(dy, rounding)
    cout, cin = dy.shape[3], x.shape[3]
    dW = np.zeros((cout, 3, 3, cin))
    d2 = dyr.reshape(-1, cout)
    for ky in range(3):
        for kx in range(3):
            xs = np.roll(xr, shift=(1 - ky, 1 - kx), axis=(1, 2)).reshape(-1, cin)
            dW[:, ky, kx, :] = d2.T @ xs
    db = np.asarray(dy, dtype=np.float64).reshape(-1, cout).sum(axis=0)
    return dW, db


def forward(theta, x, n_blocks=6, ch=32, rounding="none"):
    """x: [B,H,W,2] -> y [B,H,W], cache for backward."""
    P = unpack(np.asarray(theta, dtype=np.float64), n_blocks, ch, x.shape[3])
    x = np.asarray(x, dtype=np.float64)
    h = np.maximum(conv3x3(x, *P[0], rounding=rounding), 0.0)
    cache = {"x": x, "h": [h], "u": []}
    for k in range(n_blocks):
        Wa, ba = P[1 + 2 * k]
        Wb, bb = P[2 + 2 * k]
        u = np.maximum(conv3x3(h, Wa, ba, rounding=rounding), 0.0)
        h = h + conv3x3(u, Wb, bb, rounding=rounding)
        cache["u"].append(u)
        cache["h"].append(h)
    Wh, bh = P[-1]
    y = x[..., 0] + conv3x3(h, Wh, bh, rounding=rounding)[..., 0]
    return y, cache


def loss_and_grad(theta, x, t, n_blocks=6, ch=32, rounding="none"):
    """Mean squared error and its gradient w.r.t. the flat parameters."""
    y, c = forward(theta, x, n_blocks, ch, rounding)
    n = y.size
    r = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(r * r) / n)
    dy = (2.0 / n) * r[..., None]
    P = unpack(np.asarray(theta, dtype=np.float64), n_blocks, ch, x.shape[3])
    grads = [None] * len(P)
    h = c["h"][-1]
    dW, db = conv3x3_grad_weight(h, dy, rounding)
    grads[-1] = (dW, db)
    dh = conv3x3_grad_input(dy, P[-1][0], rounding)
    for k in reversed(range(n_blocks)):
        Wa, _ = P[1 + 2 * k]
        Wb, _ = P[2 + 2 * k]
        u, hin = c["u"][k], c["h"][k]
        grads[2 + 2 * k] = conv3x3_grad_weight(u, dh, rounding)
        du = conv3x3_grad_input(dh, Wb, rounding) * (u > 0)
        grads[1 + 2 * k] = conv3x3_grad_weight(hin, du, rounding)
        dh = dh + conv3x3_grad_input(du, Wa, rounding)
    dz0 = dh * (c["h"][0] > 0)
    grads[0] = conv3x3_grad_weight(c["x"], dz0, rounding)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return loss, flat


def adam_step(theta, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in float64."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
