"""GPU parity on the benchmark workloads themselves, against the CPU oracle:
the device assembly bit for bit against the reference's golden operator,
the uneven-extent production grid through its replayed HIP graph, the
production grid across its Kp switch with LU caching off and on, and the
graph-replay key against scratch that moves or extents that change.

Bars: assembly bit-exact; PSD after local diffusion within 1e-12 normwise
per (P,R) node (or per trimmed node slice); untouched pitch-angle columns
bit-exact.
"""
import ctypes
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import verb4d
from parity import bit_equal, node_rel_err, rel_err
from verb4d import workload as W

pytestmark = pytest.mark.gpu

LOCAL = ["local_c0_node.npz", "local_c0_rich_substeps.npz", "local_c0_neumann.npz", "local_c1_node.npz"]
ORACLE_THREADS = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("name", LOCAL)
def test_device_assembly_bit_exact_vs_reference(ctx, golden_dir, name):
    """verb_local_assembly on one node against the reference's own
    CalculationMatrix diagonals, B and C (tests/golden, written by the
    reference's Diffusion_2D assembly)."""
    import torch
    say("assembly", name)
    d = np.load(os.path.join(golden_dir, name))
    nx, ny = d["psd"].shape
    keep = {k: _dev(d[k]) for k in ("psd", "x", "y", "xl", "xu", "yl", "yu", "Dxx", "Dyy", "Dxy", "G", "S", "L")}
    g = verb4d.VerbGrid()
    g.nP, g.nR, g.nV, g.nK = 1, 1, nx, ny
    g.Vg, g.Kg = keep["x"].data_ptr(), keep["y"].data_ptr()
    g.Vl, g.Vu, g.Kl, g.Ku = (keep[k].data_ptr() for k in ("xl", "xu", "yl", "yu"))
    g.DVV, g.DKK, g.DVK = (keep[k].data_ptr() for k in ("Dxx", "Dyy", "Dxy"))
    g.G_local, g.sources, g.losses_local = (keep[k].data_ptr() for k in ("G", "S", "L"))
    g.Vl_type, g.Vu_type, g.Kl_type, g.Ku_type = (int(t) for t in d["types"])
    g.dt, g.sub_dt = float(d["dt"]), float(d["sub_dt"])
    g.local_sources, g.radial_sources, g.include_boundary = 1.0, 0.0, 1
    n = nx * ny
    diag = torch.full((9, n), np.nan, dtype=torch.float64, device="cuda")
    b0 = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    c0 = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.check(verb4d.lib().verb_local_assembly(ctx.handle, ctypes.byref(g), keep["psd"].data_ptr(),
                                               diag.data_ptr(), b0.data_ptr(), c0.data_ptr()))
    ctx.check_status()
    assert bit_equal(diag.cpu().numpy(), d["diag"])
    assert bit_equal(b0.cpu().numpy(), d["B0"])
    assert bit_equal(c0.cpu().numpy(), d["C0"])


def test_device_assembly_full_grid_vs_oracle(ctx):
    """Every node of a config-0 grid (rich inputs: sources, losses) against
    the oracle's assembly."""
    from verb4d.device import DeviceGrid
    w = W.config0(seed=61, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    diag, b0, c0 = dg.assembly()
    for iP in (0, 7, 23):
        for iR in range(w.nR):
            o = (iP, iR)
            dref, bref, cref = O.assemble_2d(w.f["Vg"][o], w.f["Kg"][o], w.nV, w.nK, w.bc["Vl"][o], w.bc["Vu"][o],
                                             w.bc["Kl"][o], w.bc["Ku"][o], (1, 1, 1, 2), w.f["DVV"][o], w.f["DKK"][o],
                                             w.f["DVK"][o], w.f["DVK"][o], w.f["G_local"][o], w.f["sources"][o],
                                             w.f["losses_local"][o], w.dt)
            node = iP * w.nR + iR
            assert bit_equal(diag[node], dref), o
            assert bit_equal(b0[node], bref) and bit_equal(c0[node], cref), o


def _oracle_local_trimmed(w, psd, k0):
    """The reference's Diffusion_2D on every node's slice [V][k0:K] (the
    trimmed y_size of src/Diffusion_2D.h:100-110), in a thread pool."""
    out = np.array(psd, copy=True)

    def node(ij):
        iP, iR = ij
        k = int(k0[iP, iR])
        o = (iP, iR)
        sl = (iP, iR, slice(None), slice(k, None))
        out[sl] = O.diffusion_2d(psd[sl], w.f["Vg"][sl], w.f["Kg"][sl], w.nV, w.nK - k, w.bc["Vl"][o][k:],
                                 w.bc["Vu"][o][k:], w.bc["Kl"][o], w.bc["Ku"][o], (1, 1, 1, 2), w.f["DVV"][sl],
                                 w.f["DKK"][sl], w.f["DVK"][sl], w.f["DVK"][sl], w.f["G_local"][sl],
                                 w.f["sources"][sl], w.f["losses_local"][sl], w.dt, w.sub_dt)

    with ThreadPoolExecutor(ORACLE_THREADS) as ex:
        list(ex.map(node, [(p, r) for p in range(w.nP) for r in range(w.nR)]))
    return out


def test_config2_uneven_graph_replay_every_node_vs_oracle(ctx):
    """The uneven-extent production workload (29x24x101x91, loss-cone plus
    +-10 drift-band extents, Kp 2 / 6 across MLT) as the bench runs it: the
    persistent 8x14 LU inside a replayed HIP graph. Every one of the 696
    nodes is compared with the oracle on its trimmed slice."""
    from verb4d.device import DeviceGrid
    say("config2: capture + replay")
    w, k0 = W.config2()
    dg = DeviceGrid(w, ctx=ctx)
    dg.set_tuning(convection_tq=1, lu_rb=8, lu_cb=14, lu_persistent=1)
    dg.step_graph(k0)  # eager run + capture
    start = dg.psd_host()
    dg.step_graph(k0)  # replay
    dg.check_status()
    got = dg.psd_host()
    say("config2: oracle convection + radial")
    cr = O.step(w, start, ORACLE_THREADS, 1, 1, 0)
    say("config2: oracle local, 696 trimmed nodes")
    ref = _oracle_local_trimmed(w, cr, k0)
    worst = 0.0
    for iP in range(w.nP):
        for iR in range(w.nR):
            k = int(k0[iP, iR])
            e = rel_err(got[iP, iR, :, k:], ref[iP, iR, :, k:])
            worst = max(worst, e)
            assert e < 1e-12, (iP, iR, k, e)
            assert bit_equal(got[iP, iR, :, :k], cr[iP, iR, :, :k]), (iP, iR, k)
    say("config2: worst node error %.3e, extents %d..%d unknowns" % (worst, w.nV * (w.nK - k0.max()),
                                                                    w.nV * (w.nK - k0.min())))


@pytest.fixture(scope="module")
def config1_switch():
    """Production grid after steps 0..48 (Kp = 2) on the GPU, and the oracle's
    steps 49 (Kp 2), 50 and 51 (Kp 6) from that state."""
    from verb4d.device import DeviceGrid
    import torch
    assert torch.cuda.is_available()
    c = verb4d.Context()
    w = W.config1(seed=1001)
    dg = DeviceGrid(w, ctx=c)
    say("config1: 49 GPU steps at Kp 2")
    for it in range(49):
        dg.step()
    dg.check_status()
    start = dg.psd_host()
    refs = []
    s = start
    for it in (49, 50, 51):
        w.set_kp(W.kp_schedule_config1(it))
        say("config1: oracle step %d (Kp %g)" % (it, w.kp))
        s = O.step(w, s, ORACLE_THREADS)
        refs.append(s)
    w.set_kp(2.0)
    return w, dg, start, refs


@pytest.mark.parametrize("caching", [False, True])
def test_config1_kp_switch_steps_vs_oracle(config1_switch, caching):
    """Steps 49 -> 51 across the Kp 2 -> 6 switch. With LU caching on, the
    factors are recomputed when the coefficients change (step 49 forces the
    first factorisation; step 50 follows the Kp update) and reused on step 51,
    as the reference's recompute flag (src/VERB4D_Solver.cpp:1091-1096)."""
    w, dg, start, refs = config1_switch
    if caching and dg.lu_cache is None:
        dg.alloc_lu_cache()
    dg.set_psd(start)
    fields = ["DKK", "DVV", "DVK", "DLL", "VP", "VR"]
    for i, it in enumerate((49, 50, 51)):
        kp = W.kp_schedule_config1(it)
        changed = (it == 49) or kp != w.kp
        if kp != w.kp:
            w.set_kp(kp)
            dg.refresh(fields)
        dg.step(use_cache=caching, recompute_lu=changed)
        dg.check_status()
        err = node_rel_err(dg.psd_host(), refs[i])
        say("config1 step %d caching=%s recompute=%s: %.3e" % (it, caching, changed, err))
        assert err < 1e-12, (it, err)
    w.set_kp(2.0)
    dg.refresh(fields)


def _uneven_k0(w, seed):
    rng = np.random.default_rng(seed)
    base = np.round(np.linspace(0.55, 0.0, w.nR) * (w.nK - 3)).astype(int)
    k0 = base[None, :] + rng.integers(-3, 4, size=(w.nP, w.nR))
    return np.clip(k0, 0, w.nK - 3).astype(np.int32)


def test_graph_replay_recaptures_after_extents_change(ctx):
    """Capture the step for extents A, run an eager step with extents B
    (which re-uploads the device extent tables), replay A: the result must
    be that of an eager step with A."""
    from verb4d.device import DeviceGrid
    w = W.config0(seed=71, rich=True)
    dg = DeviceGrid(w, ctx=ctx)
    ka, kb = _uneven_k0(w, 5), _uneven_k0(w, 6)
    assert not np.array_equal(ka, kb)
    dg.step_graph(ka)
    s1 = dg.psd_host()
    dg.step_uneven(kb)
    dg.set_psd(s1)
    dg.step_graph(ka)
    got = dg.psd_host()
    dg.set_psd(s1)
    dg.step_uneven(ka)
    assert bit_equal(got, dg.psd_host())


def test_graph_replay_recaptures_after_scratch_moves():
    """Capture the plain step, run a larger grid on the same context (its
    workspace outgrows and replaces the captured buffers), replay the first
    grid: the result must be that of an eager step."""
    from verb4d.device import DeviceGrid
    c = verb4d.Context()
    w = W.config0(seed=72)
    dg = DeviceGrid(w, ctx=c)
    dg.step_graph()
    s1 = dg.psd_host()
    big = DeviceGrid(W.Workload(nP=2, nR=3, nV=61, nK=51, seed=73, kp=2.0), ctx=c)
    big.step()
    big.check_status()
    dg.set_psd(s1)
    dg.step_graph()
    got = dg.psd_host()
    dg.set_psd(s1)
    dg.step()
    assert bit_equal(got, dg.psd_host())
