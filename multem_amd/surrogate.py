"""ctypes binding of the surrogate network entries (include/multem_surrogate.h).

The network runs in libmultem_hip.so; this module only moves host arrays in
and out.  A missing library is an error, as for the engine.
"""
import ctypes as C

import numpy as np

from . import engine

_bound = False


def _lib():
    global _bound
    L = engine.lib()
    if not _bound:
        P = C.POINTER
        fp, dp = P(C.c_float), P(C.c_double)
        L.mt_hip_sg_create.argtypes = [C.c_int, C.c_int, C.c_int, P(C.c_void_p)]
        L.mt_hip_sg_destroy.argtypes = [C.c_void_p]
        L.mt_hip_sg_param_count.argtypes = [C.c_void_p]
        L.mt_hip_sg_param_count.restype = C.c_size_t
        L.mt_hip_sg_set_params.argtypes = [C.c_void_p, fp, C.c_size_t]
        L.mt_hip_sg_get_params.argtypes = [C.c_void_p, fp, C.c_size_t]
        L.mt_hip_sg_forward.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, C.c_int, fp]
        L.mt_hip_sg_loss_grad.argtypes = [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_int, dp, fp]
        L.mt_hip_sg_train.argtypes = [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                      C.c_ulonglong, fp]
        L.mt_hip_sg_last_forward_ms.argtypes = [C.c_void_p]
        L.mt_hip_sg_last_forward_ms.restype = C.c_double
        _bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class SurrogateNet:
    """Residual CNN of n_blocks blocks of `channels` channels on GPU `device`."""

    def __init__(self, n_blocks=6, channels=32, device=0):
        self.n_blocks, self.channels = n_blocks, channels
        h = C.c_void_p()
        engine._check(_lib().mt_hip_sg_create(n_blocks, channels, device, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _lib().mt_hip_sg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self):
        return int(_lib().mt_hip_sg_param_count(self._h))

    def set_params(self, theta):
        t = _f32(theta)
        engine._check(_lib().mt_hip_sg_set_params(self._h, _fp(t), t.size))

    def get_params(self):
        t = np.zeros(self.n_params, np.float32)
        engine._check(_lib().mt_hip_sg_get_params(self._h, _fp(t), t.size))
        return t

    def forward(self, x):
        """x: [B,H,W,2] -> y [B,H,W]."""
        x = _f32(x)
        B, H, W, _ = x.shape
        y = np.zeros((B, H, W), np.float32)
        engine._check(_lib().mt_hip_sg_forward(self._h, _fp(x), B, H, W, _fp(y)))
        return y

    def last_forward_ms(self):
        return float(_lib().mt_hip_sg_last_forward_ms(self._h))

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        B, H, W, _ = x.shape
        loss = C.c_double(0)
        g = np.zeros(self.n_params, np.float32)
        engine._check(_lib().mt_hip_sg_loss_grad(self._h, _fp(x), _fp(t), B, H, W, C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, x, t, batch, steps, lr, seed=0):
        """Adam on the resident set (x [N,H,W,2], t [N,H,W]); returns per-step losses."""
        x, t = _f32(x), _f32(t)
        N, H, W, _ = x.shape
        losses = np.zeros(steps, np.float32)
        engine._check(_lib().mt_hip_sg_train(self._h, _fp(x), _fp(t), N, H, W, batch, steps, float(lr),
                                             int(seed), _fp(losses)))
        return losses


def init_params(seed, n_blocks=6, channels=32, cin=2):
    """He-uniform initial weights (flat layout of include/multem_surrogate.h);
    the second convolution of each block and the head start 10x smaller, so
    the initial network stays close to returning input channel 0."""
    rng = np.random.default_rng(seed)
    shapes = [(channels, cin)] + [(channels, channels)] * (2 * n_blocks) + [(1, channels)]
    parts = []
    for li, (co, ci) in enumerate(shapes):
        lim = np.sqrt(6.0 / (9 * ci))
        if li == len(shapes) - 1 or (li > 0 and li % 2 == 0):
            lim *= 0.1
        parts.append(rng.uniform(-lim, lim, size=co * 9 * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def normalise(adf_small, vmap):
    """Network input of one or more images: ([..., H, W] each) -> ([..., H, W, 2], scale).
    The ADF is divided by its mean (the scale the prediction is multiplied
    back by), the potential map by its own mean."""
    adf_small = np.asarray(adf_small, dtype=np.float64)
    vmap = np.asarray(vmap, dtype=np.float64)
    s = adf_small.mean(axis=(-2, -1), keepdims=True)
    v = vmap / vmap.mean(axis=(-2, -1), keepdims=True)
    return np.stack([adf_small / s, v], axis=-1).astype(np.float32), s


class Surrogate:
    """The trained network: predict(adf_small, vmap) -> many-configuration ADF image."""

    def __init__(self, theta, n_blocks=6, channels=32, device=0):
        self.net = SurrogateNet(n_blocks, channels, device)
        self.net.set_params(theta)

    @classmethod
    def load(cls, path, device=0):
        d = np.load(path, allow_pickle=False)
        return cls(d["theta"], int(d["n_blocks"]), int(d["channels"]), device)

    def predict(self, adf_small, vmap):
        x, s = normalise(adf_small, vmap)
        single = x.ndim == 3
        y = self.net.forward(x[None] if single else x)
        y = y.astype(np.float64) * (s if not single else s[..., 0, 0])
        return y[0] if single else y


def _rel_rms(pred, ref):
    pred, ref = np.asarray(pred, np.float64), np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((pred - ref) ** 2) / np.mean(ref ** 2)))


def evaluate(model, n_specimens=4, seed=2024, ns=32, n_small=2, n_big=64, n_ref=0, tilt_max=0.0):
    """Surrogate against plain multislice on held-out random specimens.

    GPU time: plain = the n_big-configuration run; surrogate = the
    n_small-configuration run + the potential-sum build + the network
    forward (device times from HIP events).  Errors are relative RMS
    against the n_big-configuration image, for the surrogate and for the
    n_small-configuration image it starts from.

    With n_ref > 0 they are also measured against an n_ref-configuration
    image drawn with another phonon seed, so that its configurations are
    independent of the other two runs', for those two and for the
    n_big-configuration image itself.  The reference's own phonon noise
    then adds to every squared error alike: E|x - ref|^2 = E|x - conv|^2 +
    v_ref for anything x independent of it, with conv the converged image;
    since the n_big and n_ref images are both configuration means,
    v_ref = E|big - ref|^2 n_big / (n_big + n_ref).  The *_conv_est means
    are sqrt(mean(e_ref^2) - v_ref) over the specimens: the errors against
    the converged image, estimated.  tilt_max: largest illumination tilt of
    the specimens (degrees)."""
    from . import engine, surrogate_data
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n_specimens):
        p, meta = surrogate_data.random_specimen(rng, ns, tilt_max=tilt_max)
        q = p.copy()
        q.pn_nconf = n_big
        big = engine.run_multislice(q).image_tot[0, 0].copy()
        t_big = engine.last_timing()["ms_total"]
        q = p.copy()
        q.pn_nconf = n_small
        small = engine.run_multislice(q).image_tot[0, 0].copy()
        t_small = engine.last_timing()["ms_total"]
        vmap = surrogate_data.potential_map(p)
        t_v = engine.last_timing()["ms_total"]
        pred = model.predict(small, vmap)
        t_f = model.net.last_forward_ms()
        row = dict(desc=meta["desc"], thick=meta["thick"], theta=p.theta, ms_plain=t_big,
                   ms_surrogate=t_small + t_v + t_f, ms_small=t_small, ms_vmap=t_v, ms_forward=t_f,
                   rel_rms_surrogate=_rel_rms(pred, big), rel_rms_small=_rel_rms(small, big))
        if n_ref:
            q = p.copy()
            q.pn_nconf = n_ref
            q.pn_seed = (p.pn_seed + 104729) % (2 ** 31 - 1) or 1   # independent configurations
            ref = engine.run_multislice(q).image_tot[0, 0].copy()
            row.update(rel_rms_surrogate_ref=_rel_rms(pred, ref), rel_rms_small_ref=_rel_rms(small, ref),
                       rel_rms_big_ref=_rel_rms(big, ref))
        rows.append(row)
    ms_p = sum(r["ms_plain"] for r in rows)
    ms_s = sum(r["ms_surrogate"] for r in rows)
    out = dict(specimens=rows, gpu_time_ratio=ms_p / ms_s,
               rel_rms_surrogate_mean=float(np.mean([r["rel_rms_surrogate"] for r in rows])),
               rel_rms_small_mean=float(np.mean([r["rel_rms_small"] for r in rows])))
    if n_ref:
        ms2 = {k: float(np.mean([r["rel_rms_%s_ref" % k] ** 2 for r in rows])) for k in ("surrogate", "small", "big")}
        v_ref = ms2["big"] * n_big / (n_big + n_ref)
        for k in ("surrogate", "small", "big"):
            out["rel_rms_%s_ref_mean" % k] = float(np.mean([r["rel_rms_%s_ref" % k] for r in rows]))
            out["rel_rms_%s_conv_est" % k] = float(np.sqrt(max(ms2[k] - v_ref, 0.0)))
    return out
