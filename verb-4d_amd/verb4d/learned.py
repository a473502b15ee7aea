"""Host side of the learned warm start (include/verb4d_hip.h, verb_unet_*):
the network handle, its parameter vector, features, forward, gradients,
training steps and the warm-started local solve. Device memory comes from
torch; every computation runs in the library's hand-written kernels.
"""
import ctypes
import os

import numpy as np

from . import GmresParams, GmresStats, lib

# 3x3 convolutions: (input channels from src0, upsampled, from src1, output
# channels, resolution level), as the library defines them
SPEC = [(16, 0, 0, 32, 0), (32, 0, 0, 32, 0), (32, 0, 0, 64, 1), (64, 0, 0, 96, 2), (96, 0, 0, 128, 3),
        (128, 1, 96, 96, 2), (96, 1, 64, 64, 1), (64, 1, 32, 32, 0)]
LIVE_IN = 5  # feature channels that carry data (of 16)
HEAD = 32
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights", "warmstart_unet.npz")


def layout():
    """[(name, offset, shape)] of the FP32 parameter vector."""
    out, o = [], 0
    for i, (c0, _, c1, co, _) in enumerate(SPEC):
        ci = c0 + c1
        out.append(("w%d" % i, o, (9, ci, co)))
        o += 9 * ci * co
        out.append(("b%d" % i, o, (co,)))
        o += co
    out.append(("wh", o, (HEAD,)))
    o += HEAD
    out.append(("bh", o, (1,)))
    o += 1
    return out, o


def unpack(params):
    """Parameter vector -> {name: array view}."""
    lay, _ = layout()
    return {n: params[o:o + int(np.prod(s))].reshape(s) for n, o, s in lay}


def init_params(seed=0):
    """He-normal 3x3 weights over the live input channels (the input layer's
    zero channels 5..15 keep zero weights), zero biases, and a small head so
    that the first delta is close to 0 (x0 close to PSD_in)."""
    rng = np.random.default_rng(seed)
    _, n = layout()
    p = np.zeros(n, dtype=np.float32)
    v = unpack(p)
    for i, (c0, _, c1, co, _) in enumerate(SPEC):
        live = (LIVE_IN if i == 0 else c0) + c1
        w = rng.standard_normal((9, c0 + c1, co)).astype(np.float32) * np.float32(np.sqrt(2.0 / (9 * live)))
        if i == 0:
            w[:, LIVE_IN:, :] = 0.0
        v["w%d" % i][...] = w
    v["wh"][...] = rng.standard_normal(HEAD).astype(np.float32) * np.float32(1e-3)
    return p


def save(path, params, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=np.asarray(params, dtype=np.float32),
                        **{k: np.asarray(v) for k, v in meta.items()})


def load(path=WEIGHTS):
    d = np.load(path)
    return d["params"].astype(np.float32), {k: d[k] for k in d.files if k != "params"}


def padded(nV, nK):
    return (nV + 7) // 8 * 8, (nK + 7) // 8 * 8


class UNet:
    """A verb_unet on a context: batches of up to max_batch nodes of
    height x width (multiples of 8); train=True adds gradients and Adam."""

    def __init__(self, ctx, max_batch, height, width, train=False, params=None):
        self.ctx = ctx
        self.max_batch, self.H, self.W = int(max_batch), int(height), int(width)
        h = ctypes.c_void_p()
        ctx.check(lib().verb_unet_create(ctx.handle, self.max_batch, self.H, self.W, 1 if train else 0,
                                         ctypes.byref(h)))
        self.handle = h
        self.n_params = int(lib().verb_unet_param_count())
        self.set_params(init_params() if params is None else params)

    def close(self):
        if getattr(self, "handle", None):
            lib().verb_unet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_params
        self.ctx.check(lib().verb_unet_set_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self.ctx.check(lib().verb_unet_get_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def features(self, dg, node0, count, psd=None):
        """(feat int16 [count][Hp][Wp][16], pin f32 [count][Hp][Wp], scale
        f64 [count]) device tensors for nodes node0.. of a DeviceGrid."""
        import torch
        Hp, Wp = padded(dg.w.nV, dg.w.nK)
        feat = torch.empty((count, Hp, Wp, 16), dtype=torch.int16, device="cuda")
        pin = torch.empty((count, Hp, Wp), dtype=torch.float32, device="cuda")
        scale = torch.empty((count,), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        x = dg.psd if psd is None else psd
        self.ctx.check(lib().verb_unet_features(self.handle, ctypes.byref(dg.grid), x.data_ptr(), int(node0),
                                                int(count), feat.data_ptr(), pin.data_ptr(), scale.data_ptr()))
        # the library's stream wrote them: finish before torch reads them
        self.ctx.synchronize()
        return feat, pin, scale

    def normalize(self, dg, node0, count, scale, psd=None):
        import torch
        Hp, Wp = padded(dg.w.nV, dg.w.nK)
        out = torch.empty((count, Hp, Wp), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        x = dg.psd if psd is None else psd
        self.ctx.check(lib().verb_unet_normalize(self.handle, ctypes.byref(dg.grid), x.data_ptr(), int(node0),
                                                 int(count), scale.data_ptr(), out.data_ptr()))
        self.ctx.synchronize()
        return out

    def forward(self, feat):
        import torch
        b, h, w, _ = feat.shape
        delta = torch.empty((b, h, w), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()  # feat may come from torch's stream
        self.ctx.check(lib().verb_unet_forward(self.handle, feat.data_ptr(), b, h, w, delta.data_ptr()))
        self.ctx.synchronize()
        return delta

    def gradients(self, feat, pin, pout):
        """(gradient vector (host f32), loss) of a batch."""
        import torch
        b, h, w, _ = feat.shape
        g = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        loss = ctypes.c_float()
        torch.cuda.synchronize()
        self.ctx.check(lib().verb_unet_gradients(self.handle, feat.data_ptr(), pin.data_ptr(), pout.data_ptr(), b, h,
                                                 w, g.data_ptr(), ctypes.byref(loss)))
        self.ctx.synchronize()
        return g.cpu().numpy(), float(loss.value)

    def train_step(self, feat, pin, pout, lr, want_loss=False):
        """One Adam step on a batch. The batch tensors may come straight from
        torch operations: torch's stream is finished before the library
        reads them, and the library's before they can be freed."""
        import torch
        b, h, w, _ = feat.shape
        loss = ctypes.c_float()
        torch.cuda.current_stream().synchronize()
        self.ctx.check(lib().verb_unet_train_step(self.handle, feat.data_ptr(), pin.data_ptr(), pout.data_ptr(), b, h,
                                                  w, ctypes.c_float(lr), ctypes.byref(loss) if want_loss else None))
        self.ctx.synchronize()
        return float(loss.value) if want_loss else None


def local_learned(dg, net, params=None, stats=True):
    """verb_local_diffusion_learned on a DeviceGrid (stats dict or None)."""
    from .device import DeviceGrid
    p = params or DeviceGrid.gmres_params_capped()
    st = GmresStats() if stats else None
    dg.ctx.check(lib().verb_local_diffusion_learned(dg.ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(),
                                                    net.handle, ctypes.byref(p), ctypes.byref(st) if st else None))
    return st.as_dict() if st else None


def step_learned(dg, net, params=None, stats=False):
    from .device import DeviceGrid
    p = params or DeviceGrid.gmres_params_capped()
    st = GmresStats() if stats else None
    dg.ctx.check(lib().verb_step_learned(dg.ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(), net.handle,
                                         ctypes.byref(p), ctypes.byref(st) if st else None))
    return st.as_dict() if st else None


__all__ = ["UNet", "SPEC", "layout", "unpack", "init_params", "save", "load", "padded", "local_learned",
           "step_learned", "WEIGHTS", "GmresParams"]
