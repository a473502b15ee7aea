"""The CPU reference of the warm-start U-Net (tests/unet_reference.py) is
itself pinned: its autograd gradients against central finite differences of
its loss (float64, no BF16 rounding) on a tiny input, and its BF16
emulation against the bit-level round-to-nearest-even rule."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import unet_reference as R  # noqa: E402
from verb4d import learned as L  # noqa: E402


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-30, 65504.0], dtype=np.float32)
    bits = R.f32_to_bf16_bits(x)
    back = R.bf16_bits_to_f32(bits)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(back, ref)
    # a tie rounds to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert back[1] == 1.0


def test_parameter_layout_matches_spec():
    lay, n = L.layout()
    assert n == 512065
    assert [s for name, _, s in lay if name.startswith("w") and name != "wh"] == [
        (9, ci, co) for ci, co in R.torch_layout_check()]


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(4)
    p = L.init_params(seed=5).astype(np.float64)
    v = L.unpack(p)
    for k in v:
        if k.startswith("b"):
            v[k][...] = 0.05 * rng.standard_normal(v[k].shape)
    v["wh"][...] = 0.1 * rng.standard_normal(32)
    feat = np.zeros((2, 16, 8, 16))
    feat[..., :5] = rng.standard_normal((2, 16, 8, 5))
    pin = rng.random((2, 16, 8))
    pout = pin * (1.0 + 0.1 * rng.standard_normal((2, 16, 8)))
    loss, grads = R.loss_and_grads(p, feat, pin, pout, exact=True)
    lay, _ = L.layout()
    checked = 0
    for name, off, shape in lay:
        size = int(np.prod(shape))
        g = grads[name].ravel()
        for idx in rng.choice(size, size=min(3, size), replace=False):
            if name == "w0" and (idx // 32) % 16 >= 5:
                continue
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[off + idx] += e
            pm[off + idx] -= e
            fd = (R.loss_and_grads(pp, feat, pin, pout, exact=True)[0]
                  - R.loss_and_grads(pm, feat, pin, pout, exact=True)[0]) / (2 * e)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7 * abs(loss), (name, idx, fd, g[idx])
            checked += 1
    assert checked >= 40
