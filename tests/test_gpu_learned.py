"""The learned warm start on the GPU: its hand-written kernels (features,
forward, backward) against the CPU FP32 reference (tests/unet_reference.py,
torch on the CPU with the same BF16 rounding of activations), training
steps, and the warm-started local solve against the oracle.

Bars: features within one BF16 ulp of the numpy formula; forward delta and
every parameter gradient within 2e-2 / 5e-2 relative L2 (FP32 summation
order and BF16 rounding of intermediate values differ); the learned local
solve within 1e-8 normwise per node of the oracle.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import unet_reference as R
import verb4d
from parity import node_rel_err
from verb4d import learned as L
from verb4d import workload as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


@pytest.fixture(scope="module")
def small(ctx):
    """A config-0 grid (31x31 nodes -> 32x32 padded) after one plain step:
    features of 6 nodes and their plain-solver targets."""
    from verb4d.device import DeviceGrid
    w = W.config0(seed=91)
    dg = DeviceGrid(w, ctx=ctx)
    dg.step(1, 1, 0)
    net = L.UNet(ctx, 8, 32, 32, train=True, params=L.init_params(seed=3))
    feat, pin, scale = net.features(dg, 30, 6)  # (synchronises the library's stream)
    psd_in = dg.psd.clone()
    dg.step(0, 0, 1)
    pout = net.normalize(dg, 30, 6, scale)
    ctx.synchronize()
    return w, dg, net, feat, pin, scale, pout, psd_in


def test_features_match_formula(small):
    import torch
    w, dg, net, feat, pin, scale, pout, psd_in = small
    f = R.bf16_bits_to_f32(feat.cpu().numpy().view(np.uint16))
    x = psd_in.cpu().numpy()
    for j in range(6):
        iP, iR = divmod(30 + j, w.nR)
        o = (iP, iR)
        s = np.abs(x[o]).max()
        daa, dvv, dvk, e = w.f["DKK"][o], w.f["DVV"][o], w.f["DVK"][o], w.f["Vg"][o]
        ch = [(np.log10(np.maximum(daa, 1e-30)) + 5) / 3, (np.log10(np.maximum(dvv / e ** 2, 1e-30)) + 5) / 3,
              dvk / np.sqrt(np.maximum(dvv * daa, 1e-300)), np.log10(np.maximum(x[o] / s, 1e-30)) / 15 + 1,
              np.full_like(daa, w.dt / 600.0)]
        for c in range(5):
            want = R.bf16_bits_to_f32(R.f32_to_bf16_bits(ch[c].astype(np.float32)))
            got = f[j, :w.nV, :w.nK, c]
            ulp = np.abs(want) * 2.0 ** -7 + 1e-30
            assert np.all(np.abs(got - want) <= ulp), (j, c)
        assert np.all(f[j, :, :, 5:] == 0) and np.all(f[j, w.nV:, :, :] == 0) and np.all(f[j, :, w.nK:, :] == 0)
        assert abs(float(scale[j]) - s) == 0.0
        assert np.allclose(pin[j, :w.nV, :w.nK].cpu().numpy(), (x[o] / s).astype(np.float32), rtol=0, atol=0)
    assert float(torch.abs(pin[:, w.nV:, :]).max()) == 0.0


def test_forward_matches_cpu_reference(small):
    w, dg, net, feat, pin, scale, pout, psd_in = small
    params = L.init_params(seed=3)
    v = L.unpack(params)
    rng = np.random.default_rng(0)
    for k in v:
        if k.startswith("b"):
            v[k][...] = 0.05 * rng.standard_normal(v[k].shape)
    v["wh"][...] = 0.2 * rng.standard_normal(32)
    net.set_params(params)
    got = net.forward(feat).cpu().numpy()
    f = R.bf16_bits_to_f32(feat.cpu().numpy().view(np.uint16))
    ref, _ = R.forward(params, f)
    e = rel_l2(got, ref.detach().numpy())
    say("forward rel L2 %.2e" % e)
    assert e < 2e-2


def test_gradients_match_cpu_reference(small):
    w, dg, net, feat, pin, scale, pout, psd_in = small
    params = L.init_params(seed=4)
    v = L.unpack(params)
    rng = np.random.default_rng(1)
    v["wh"][...] = 0.2 * rng.standard_normal(32)
    net.set_params(params)
    g, loss = net.gradients(feat, pin, pout)
    f = R.bf16_bits_to_f32(feat.cpu().numpy().view(np.uint16))
    rloss, rg = R.loss_and_grads(params, f, pin.cpu().numpy(), pout.cpu().numpy())
    say("loss gpu %.6e cpu %.6e" % (loss, rloss))
    assert abs(loss - rloss) <= 1e-2 * abs(rloss)
    gv = L.unpack(g)
    worst = 0.0
    for k in rg:
        a, b = gv[k], rg[k]
        if k == "w0":
            a, b = a[:, :5, :], b[:, :5, :]
            assert np.all(gv[k][:, 5:, :] == 0)
        e = rel_l2(a, b)
        worst = max(worst, e)
        say("grad %-3s rel L2 %.2e (|g| %.2e)" % (k, e, np.linalg.norm(b)))
        assert e < 5e-2, k
    say("worst gradient rel L2 %.2e" % worst)


def test_training_steps_reduce_the_loss(small):
    w, dg, net, feat, pin, scale, pout, psd_in = small
    net.set_params(L.init_params(seed=5))
    first = net.train_step(feat, pin, pout, 1e-3, want_loss=True)
    for _ in range(40):
        last = net.train_step(feat, pin, pout, 1e-3, want_loss=True)
    say("loss %.4e -> %.4e" % (first, last))
    assert last < 0.7 * first


def _oracle_local(w, psd, threads=16):
    out = np.array(psd, copy=True)

    def node(ij):
        iP, iR = ij
        o = (iP, iR)
        out[o] = O.diffusion_2d(psd[o], w.f["Vg"][o], w.f["Kg"][o], w.nV, w.nK, w.bc["Vl"][o], w.bc["Vu"][o],
                                w.bc["Kl"][o], w.bc["Ku"][o], (1, 1, 1, 2), w.f["DVV"][o], w.f["DKK"][o],
                                w.f["DVK"][o], w.f["DVK"][o], w.f["G_local"][o], w.f["sources"][o],
                                w.f["losses_local"][o], w.dt, w.sub_dt)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(node, [(p, r) for p in range(w.nP) for r in range(w.nR)]))
    return out


def _trained_params():
    if os.path.exists(L.WEIGHTS):
        return L.load()[0], True
    return L.init_params(seed=0), False


def test_learned_local_solve_vs_oracle(ctx):
    """The warm-started local solve (network x0, FP32 GMRES judged by the FP64
    residual, band LU for the nodes that miss 1e-10 within 8 iterations) on
    the production grid against the oracle."""
    from verb4d.device import DeviceGrid
    w = W.config1(seed=1101)
    dg = DeviceGrid(w, ctx=ctx)
    dg.step(1, 1, 0)
    start = dg.psd_host()
    params, trained = _trained_params()
    net = L.UNet(ctx, 696, 104, 96, train=False, params=params)
    st = L.local_learned(dg, net)
    dg.check_status()
    got = dg.psd_host()
    ref = _oracle_local(w, start)
    err = node_rel_err(got, ref)
    say("learned local solve (trained=%s): %s, max node error %.2e" % (trained, st, err))
    assert st["converged"] + st["fallback"] == st["nodes"] == 696
    assert st["max_iterations"] <= 8
    assert err < 1e-8


def test_warm_start_gain_report(ctx):
    """What the trained network gains, on the production grid (seed 1102,
    not in the training data) at Kp = 2 and Kp = 6, from the incoming PSD
    and from the network's x0:
      8-iteration rule (tol 1e-10): the mean log10 of the relative FP64
        residual each node reaches, and the fallback count;
      up to 160 inner iterations (GMRES(4) cycles, so the count moves in
        steps of 4): iterations summed over the nodes, and the nodes that
        still miss tol.
    The trained network must reach a lower residual under the 8-iteration
    rule at both Kp values."""
    import torch
    from verb4d.device import DeviceGrid
    params, trained = _trained_params()
    w = W.config1(seed=1102)
    dg = DeviceGrid(w, ctx=ctx)
    net = L.UNet(ctx, 696, 104, 96, train=False, params=params)
    wide = DeviceGrid.gmres_params(tol=1e-10, restart=4, max_cycles=40, inner_tol=1e-12)
    capped = DeviceGrid.gmres_params_capped()
    rows = {}
    for kp in (2.0, 6.0):
        w.set_kp(kp)
        dg.refresh(["DKK", "DVV", "DVK", "DLL", "VP", "VR"])
        dg.set_psd(w.psd0)
        dg.step(1, 1, 0)
        dg.ctx.synchronize()
        start = dg.psd.clone()
        feat, pin, scale = net.features(dg, 0, 696)
        delta = net.forward(feat)
        x0 = (start.view(696, w.nV, w.nK) +
              scale.view(696, 1, 1) * delta[:, :w.nV, :w.nK].double()).view_as(start).contiguous()
        res = {}
        for name, guess in (("psd", None), ("net", x0)):
            for rule, prm in (("capped", capped), ("wide", wide)):
                dg.psd.copy_(start)
                torch.cuda.synchronize()
                res[(name, rule)] = dg.local_gmres(prm, x0=guess)
        rows[kp] = res
        c0, c1 = res[("psd", "capped")], res[("net", "capped")]
        w0, w1 = res[("psd", "wide")], res[("net", "wide")]
        say("Kp %g, 8 iterations: mean log10 residual psd %.2f net %.2f, fallback psd %d net %d of 696; "
            "160 iterations: iterations psd %d net %d, unconverged psd %d net %d (trained=%s)" % (
                kp, c0["mean_log10_residual"], c1["mean_log10_residual"], c0["fallback"], c1["fallback"],
                w0["iterations"], w1["iterations"], w0["fallback"], w1["fallback"], trained))
    if trained:
        for kp in (2.0, 6.0):
            assert rows[kp][("net", "capped")]["mean_log10_residual"] < rows[kp][("psd", "capped")][
                "mean_log10_residual"], kp
