"""The data assimilation's Kalman filter in the (MLT, R) plane on the GPU
(da_conv.hip, verb_da_convection) against the oracle (oracle/da_oracle.py:
conv_model_matrix, kalman_conv_plane, itself pinned bit for bit to the
reference's matrix operations) on a range of (V,K) planes of the 24 x 10 x
31 x 31 grid: velocities whose Courant numbers need one to five sub-steps,
a loss array and a saturation density, observations at some points of
every plane and at all points of one; the model matrices rebuilt on the
first call and reused on the second (which starts from the first's
covariance), in linear and in log space, with the boundary Q reset and
growing, with Q kept between the calls (linear space) or rebuilt (log
space). A plane without observations outside log space makes the
reference's solve exit; here the call returns VERB_ERR_SINGULAR.

Bars (the products' summation order differs from OpenBLAS's): the model
matrices, analysis PSD and covariances within 1e-12 of each plane's
largest value; planes outside the range and the other points untouched,
bit for bit."""
import math

import numpy as np
import pytest

import da_oracle as D
import verb4d
from verb4d import workload as W

pytestmark = pytest.mark.gpu

DAY = 86400.0


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _setup(seed):
    """config0 with velocities per day and a DA step of 0.0255 days (sub-step
    counts 1 to 5, M1d = M^5), loss = -rate * PSD with rates up to 0.02 per
    day, a saturation density at 1.5 the PSD with a few points reached."""
    w = W.config0(seed=seed)
    # per day, scaled plane by plane (0.4 to 2.4) to spread the sub-step counts
    fac = 0.4 + 2.0 * (np.arange(w.nV * w.nK) % 7).reshape(w.nV, w.nK) / 6
    w.f["VP"] = w.f["VP"] * DAY * fac
    w.f["VR"] = w.f["VR"] * DAY * fac
    rng = np.random.default_rng(seed)
    loss = -rng.uniform(0, 0.02, w.shape) * w.psd0
    satd = w.psd0 * 1.5
    hit = rng.random(w.shape) < 0.05
    satd[hit] = w.psd0[hit] * 0.9
    obs = np.where(rng.random(w.shape) < 0.3, w.psd0 * (1 + 0.2 * rng.standard_normal(w.shape)), np.nan)
    obs = np.abs(obs)
    return w, loss, satd, obs


def _plane(a, q, nK):
    return a[:, :, q // nK, q % nK]


def _oracle_planes(w, psd, loss, satd, obs, Pa, M, planes, dt, corr, me, oe, use_log, growing, recompute):
    """oracle forecast/analysis of the given planes: psd, Pa, M updated copies"""
    psd, Pa, M = psd.copy(), Pa.copy(), M.copy()
    dP = w.f["Pg"][1, 0, 0, 0] - w.f["Pg"][0, 0, 0, 0]
    dR = w.f["Rg"][0, 1, 0, 0] - w.f["Rg"][0, 0, 0, 0]
    for q in planes:
        x = _plane(psd, q, w.nK)
        if recompute:
            rate = _plane(loss, q, w.nK) / x
            sat = x >= _plane(satd, q, w.nK)
            M[q] = D.conv_model_matrix(_plane(w.f["VP"], q, w.nK), _plane(w.f["VR"], q, w.nK), rate, sat, dt, dP, dR,
                                       use_log, corr)
        xa, Pa[q], _ = D.kalman_conv_plane(x, Pa[q], M[q], _plane(obs, q, w.nK), None, None, dt, corr, me, oe,
                                           use_log, growing)
        psd[:, :, q // w.nK, q % w.nK] = xa
    return psd, Pa, M


def _plane_err(got, want, planes):
    e = 0.0
    for q in planes:
        e = max(e, float(np.abs(got[q] - want[q]).max() / max(np.abs(want[q]).max(), 1e-300)))
    return e


def _psd_err(got, want, planes, nK):
    e = 0.0
    for q in planes:
        g, o = _plane(got, q, nK), _plane(want, q, nK)
        e = max(e, float(np.abs(g - o).max() / np.abs(o).max()))
    return e


@pytest.mark.parametrize("use_log,growing", [(False, False), (True, True)])
def test_da_convection_vs_oracle(ctx, use_log, growing):
    import torch
    from verb4d.device import DeviceGrid
    w, loss, satd, obs = _setup(151)
    NQ, N = w.nV * w.nK, w.nP * w.nR
    q0, nq = (300, 24) if not use_log else (520, 16)
    planes = range(q0, q0 + nq)
    obs[:, :, q0 // w.nK, q0 % w.nK] = np.abs(w.psd0[:, :, q0 // w.nK, q0 % w.nK]) * 1.1  # observed everywhere
    dt, corr, me, oe = 0.0255, 2.0, 1e-2, 4e-2
    # the sub-step counts the range covers
    dP = w.f["Pg"][1, 0, 0, 0] - w.f["Pg"][0, 0, 0, 0]
    dR = w.f["Rg"][0, 1, 0, 0] - w.f["Rg"][0, 0, 0, 0]
    nts = {max(1, math.ceil((np.abs(_plane(w.f["VP"], q, w.nK)) * dt / dP +
                             np.abs(_plane(w.f["VR"], q, w.nK)) * dt / dR).max())) for q in planes}
    assert len(nts) >= 2, nts
    dg = DeviceGrid(w, ctx=ctx)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    tobs, tloss, tsat = dev(obs), dev(loss), dev(satd)
    cov = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
    model = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
    # the linear-space case keeps Q between the calls, the log-space one rebuilds it
    qn = None if use_log else torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
    x = w.psd0.copy()
    Pa = np.zeros((NQ, N, N))
    M = np.zeros((NQ, N, N))
    for rep in range(2):
        dg.da_convection(tobs, tloss, cov, model, dt, corr, me, oe, saturation=tsat, use_log=use_log,
                         growing_Q_at_boundary=growing, recompute_model=(rep == 0), plane_begin=q0, plane_count=nq,
                         process_noise=qn)
        dg.check_status()
        x, Pa, M = _oracle_planes(w, x, loss, satd, obs, Pa, M, planes, dt, corr, me, oe, use_log, growing,
                                  rep == 0)
        got = dg.psd_host()
        gm, gc = model.cpu().numpy(), cov.cpu().numpy()
        em = _plane_err(gm, M, planes)
        ec = _plane_err(gc, Pa, planes)
        ex = _psd_err(got, x, planes, w.nK)
        print("log %d rep %d: model %.1e covariance %.1e psd %.1e (sub-steps %s)" % (use_log, rep, em, ec, ex,
                                                                                   sorted(nts)), flush=True)
        assert em < 1e-12 and ec < 1e-12 and ex < 1e-12, (rep, em, ec, ex)
        # the planes outside the range are untouched
        flat = got.reshape(w.nP, w.nR, NQ)
        keep = np.r_[0:q0, q0 + nq:NQ]
        assert np.array_equal(flat[:, :, keep], w.psd0.reshape(w.nP, w.nR, NQ)[:, :, keep])
        assert not np.any(gc[keep]) and not np.any(gm[keep])
        if qn is not None:
            assert not np.any(qn.cpu().numpy()[keep]) and np.all(qn.cpu().numpy()[q0].diagonal() > 0)
    # the plane observed everywhere moved towards its observations
    q = q0
    before = np.abs(_plane(w.psd0, q, w.nK) - _plane(obs, q, w.nK)).sum()
    after = np.abs(_plane(dg.psd_host(), q, w.nK) - _plane(obs, q, w.nK)).sum()
    assert after < before


def test_da_convection_no_observations_singular(ctx):
    """one plane without observations outside log space: S = 0, the
    reference's trans_solve exits; the call reports VERB_ERR_SINGULAR"""
    import torch
    from verb4d.device import DeviceGrid
    w, loss, satd, obs = _setup(152)
    NQ, N = w.nV * w.nK, w.nP * w.nR
    q = 40
    obs[:, :, q // w.nK, q % w.nK] = np.nan
    dg = DeviceGrid(w, ctx=ctx)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    cov = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
    model = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
    with pytest.raises(verb4d.VerbError) as e:
        dg.da_convection(dev(obs), dev(loss), cov, model, 0.0255, 2.0, 1e-2, 4e-2, plane_begin=q, plane_count=1)
    assert e.value.status == -11
    dg.check_status()


def test_da_convection_handle_survives_graph_captures():
    """The filter's rocBLAS handle lives in the context from the first
    verb_da_convection call until verb_ctx_destroy: step graphs captured in
    between (here two, for two grids) leave it alone, so a later filter call
    gives the same result as the first one from the same inputs, and the
    context is destroyed once."""
    import torch
    from verb4d.device import DeviceGrid
    w, loss, satd, obs = _setup(153)
    NQ, N = w.nV * w.nK, w.nP * w.nR
    q0, nq = 100, 6
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    tobs, tloss, tsat = dev(obs), dev(loss), dev(satd)
    c = verb4d.Context()
    dg = DeviceGrid(w, ctx=c)
    w2 = W.config0(seed=154)
    w2.psd0 = w2.psd0 * 0.5
    dg2 = DeviceGrid(w2, ctx=c)
    results = []
    for rep in range(2):
        cov = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
        model = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
        dg.set_psd(w.psd0)
        dg.da_convection(tobs, tloss, cov, model, 0.0255, 2.0, 1e-2, 4e-2, saturation=tsat, plane_begin=q0,
                         plane_count=nq)
        dg.check_status()
        results.append((dg.psd_host(), cov.cpu().numpy()))
        if rep == 0:
            dg.step_graph()  # capture for this grid
            dg2.step_graph()  # a different grid: a second capture
            dg.check_status()
    assert np.array_equal(results[0][0], results[1][0])
    assert np.array_equal(results[0][1], results[1][1])
    c.close()
