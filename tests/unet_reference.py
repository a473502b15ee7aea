"""CPU FP32 reference of the warm-start U-Net (torch on the CPU), the checker
of the library's hand-written forward and backward kernels. The library
rounds every activation to BF16 after bias and ReLU (and after pooling);
the reference does the same with torch's round-to-nearest-even BF16
conversion, so the two differ only by FP32 summation order (and the BF16
rounding of the library's pre-activation gradients).
"""
import numpy as np
import torch
import torch.nn.functional as F

from verb4d.learned import SPEC, unpack


def bf16_bits_to_f32(bits):
    """uint16 / int16 BF16 bit patterns -> float32 values (numpy)."""
    b = np.asarray(bits).astype(np.uint16).astype(np.uint32) << 16
    return b.view(np.float32)


def f32_to_bf16_bits(x):
    """float32 -> BF16 bit patterns, round to nearest even (numpy)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).astype(np.uint16)


def forward(params, feat_nhwc, requires_grad=False, exact=False):
    """delta [B][H][W] and the parameter tensors (torch, fp32) for input
    features [B][H][W][16] (float32 values of the BF16 features). exact:
    float64 throughout and no BF16 rounding (the finite-difference check of
    the reference itself)."""
    dt = torch.float64 if exact else torch.float32

    def _bf(x):
        return x if exact else x.to(torch.bfloat16).to(torch.float32)

    v = unpack(np.asarray(params, dtype=np.float64 if exact else np.float32))
    t = {}
    for k, a in v.items():
        tt = torch.tensor(np.array(a), dtype=dt)
        if requires_grad:
            tt.requires_grad_(True)
        t[k] = tt

    def conv(x, i):
        w = t["w%d" % i]  # [9][cin][cout] -> OIHW
        wt = _bf(w).reshape(3, 3, w.shape[1], w.shape[2]).permute(3, 2, 0, 1)
        return _bf(F.relu(F.conv2d(x, wt, t["b%d" % i], padding=1)))

    x = torch.tensor(np.asarray(feat_nhwc), dtype=dt).permute(0, 3, 1, 2)
    up = lambda z: F.interpolate(z, scale_factor=2, mode="nearest")  # noqa: E731
    pool = lambda z: _bf(F.avg_pool2d(z, 2))  # noqa: E731
    a0 = conv(x, 0)
    a1 = conv(a0, 1)
    a2 = conv(pool(a1), 2)
    a3 = conv(pool(a2), 3)
    a4 = conv(pool(a3), 4)
    a5 = conv(torch.cat([up(a4), a3], 1), 5)
    a6 = conv(torch.cat([up(a5), a2], 1), 6)
    a7 = conv(torch.cat([up(a6), a1], 1), 7)
    delta = torch.einsum("bchw,c->bhw", a7, t["wh"]) + t["bh"]
    return delta, t


def loss_and_grads(params, feat_nhwc, pin, pout, exact=False):
    """The library's loss, sum (pin + delta - pout)^2 / (B H W), and the
    gradient of every parameter tensor (dict of numpy arrays)."""
    delta, t = forward(params, feat_nhwc, requires_grad=True, exact=exact)
    dt = torch.float64 if exact else torch.float32
    pin_t = torch.tensor(np.asarray(pin), dtype=dt)
    pout_t = torch.tensor(np.asarray(pout), dtype=dt)
    x0 = pin_t + delta
    loss = ((x0 - pout_t) ** 2).sum() / delta.numel()
    loss.backward()
    return float(loss.detach()), {k: v.grad.numpy().copy() for k, v in t.items()}


def torch_layout_check():
    """The parameter layout covers the spec (used by a CPU test)."""
    return [(c0 + c1, co) for c0, _, c1, co, _ in SPEC]
