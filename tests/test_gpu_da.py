"""The data assimilation's Kalman update along L on the GPU (da.hip,
verb_da_radial) against the oracle (oracle/da_oracle.py, itself pinned
bit for bit to the reference's matrix operations), on every (V,K) line of a
P slice, two assimilations in a row (the second from the first's
covariance), with observations at some points, at none and at all.

Bars (the products' summation order differs from OpenBLAS's): the model
matrices, the analysis PSD and the covariances within 1e-13 of
each line's largest value."""
import numpy as np
import pytest

import da_oracle as D
import verb4d
from verb4d import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _obs(w, rng, iP):
    obs = np.full(w.shape, np.nan)
    x = w.psd0[iP]
    o = x * (1 + 0.2 * rng.standard_normal(x.shape))
    keep = rng.random(x.shape) < 0.4
    keep[:, 0, 0] = False                      # a line with no observation
    keep[:, 1, 1] = True                       # and one observed everywhere
    obs[iP] = np.where(keep, o, np.nan)
    return obs


def _line_err(a, b):
    """max over lines (last two axes reduced) of |a - b| / max |b|"""
    num = np.abs(a - b).reshape(a.shape[0], a.shape[1], -1).max(axis=-1)
    den = np.abs(b).reshape(b.shape[0], b.shape[1], -1).max(axis=-1)
    return float((num / np.maximum(den, 1e-300)).max())


@pytest.mark.parametrize("which", ["config0", "production"])
def test_da_radial_vs_oracle(ctx, which):
    import torch
    from verb4d.device import DeviceGrid
    w = W.config0(seed=141) if which == "config0" else W.config1(seed=1401)
    iP = 3 if which == "config0" else 0
    rng = np.random.default_rng(7)
    obs = _obs(w, rng, iP)
    dg = DeviceGrid(w, ctx=ctx)
    n = w.nR
    cov = torch.zeros((w.nV, w.nK, n, n), dtype=torch.float64, device="cuda")
    Fm = torch.zeros((w.nV, w.nK, n, n), dtype=torch.float64, device="cuda")
    tobs = torch.from_numpy(obs).cuda()
    scale = float(np.abs(w.psd0[iP]).max())
    me, oe, dt = 1e-2 * scale ** 2, 4e-2 * scale ** 2, 3600.0
    x = w.psd0.copy()
    Pa = np.zeros((w.nV, w.nK, n, n))
    # the oracle's model matrices
    Fo = np.empty_like(Pa)
    for iV in range(w.nV):
        for iK in range(w.nK):
            Fo[iV, iK] = D.model_matrix(w.f["Lg"][iP, :, iV, iK], w.f["DLL"][iP, :, iV, iK],
                                        w.f["G_radial"][iP, :, iV, iK], w.types["Ll"], w.types["Lu"], dt)
    for rep in range(2):
        dg.da_radial(tobs, cov, dt, me, oe, p_index=iP, model_matrix=Fm)
        dg.check_status()
        x, Pa = D.da_radial(w, x, obs, Pa, iP, dt, me, oe)
        got = dg.psd_host()
        assert np.array_equal(np.delete(got, iP, axis=0), np.delete(x, iP, axis=0))  # other slices untouched
        ex = _line_err(np.moveaxis(got[iP], 0, -1), np.moveaxis(x[iP], 0, -1))
        ep = _line_err(cov.cpu().numpy(), Pa)
        ef = _line_err(Fm.cpu().numpy(), Fo)
        print("%s rep %d: psd %.1e covariance %.1e F %.1e" % (which, rep, ex, ep, ef), flush=True)
        assert ef < 1e-13 and ex < 1e-13 and ep < 1e-13, (rep, ex, ep, ef)
    # the line without observations kept its forecast in both updates
    assert np.array_equal(dg.psd_host()[iP, :, 0, 0], w.psd0[iP, :, 0, 0])
