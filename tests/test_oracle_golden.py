"""The CPU oracle against the golden vectors the reference's own code wrote
(tests/golden/make_golden.py). Convection, radial diffusion and the local
assembly must match bit for bit; the band solve within 1e-12 (the reference
factors with OpenBLAS's blocked dgbtrf, the oracle restates unblocked
dgbtf2, so the rounding differs)."""
import os

import numpy as np
import pytest

import oracle as O
from parity import bit_equal, node_rel_err, rel_err

LOCAL = ["local_c0_node.npz", "local_c0_rich_substeps.npz", "local_c0_neumann.npz", "local_c1_node.npz"]


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


@pytest.mark.parametrize("name", LOCAL)
def test_local_assembly_bit_exact(golden_dir, name):
    d = load(golden_dir, name)
    nx, ny = d["psd"].shape
    nsub = max(1, int(float(d["dt"]) / float(d["sub_dt"])))
    diag, B0, C0 = O.assemble_2d(d["x"], d["y"], nx, ny, d["xl"], d["xu"], d["yl"], d["yu"], tuple(d["types"]),
                                 d["Dxx"], d["Dyy"], d["Dxy"], d["Dxy"], d["G"], d["S"], d["L"],
                                 float(d["dt"]) / nsub)
    assert bit_equal(diag, d["diag"])
    assert bit_equal(B0, d["B0"])
    assert bit_equal(C0, d["C0"])


@pytest.mark.parametrize("name", LOCAL)
def test_local_diffusion(golden_dir, name):
    d = load(golden_dir, name)
    nx, ny = d["psd"].shape
    out = O.diffusion_2d(d["psd"], d["x"], d["y"], nx, ny, d["xl"], d["xu"], d["yl"], d["yu"], tuple(d["types"]),
                         d["Dxx"], d["Dyy"], d["Dxy"], d["Dxy"], d["G"], d["S"], d["L"], float(d["dt"]),
                         float(d["sub_dt"]))
    assert rel_err(out, d["out"]) < 1e-12


@pytest.mark.parametrize("name", ["radial_c0_lines.npz", "radial_c1_lines.npz"])
def test_radial_bit_exact(golden_dir, name):
    d = load(golden_dir, name)
    for i in range(len(d["psd"])):
        lt, ut = (int(t) for t in d["types"][i])
        out = O.diffusion_1d(d["psd"][i], d["x"][i], d["psd"].shape[1], float(d["lbc"][i]), float(d["ubc"][i]),
                             lt, ut, d["D"][i], d["G"][i], d["S"][i], d["L"][i], 300.0)
        assert bit_equal(out, d["out"][i]), i


def test_quickest6_bit_exact(golden_dir):
    d = load(golden_dir, "quickest6_lines.npz")
    for i in range(int(d["count"])):
        g = lambda k: d["%d_%s" % (i, k)]  # noqa: E731
        n = int(g("n"))
        out = O.quickest6(g("psd"), g("x"), n, float(g("lbc")), float(g("ubc")), int(g("lt")), int(g("ut")),
                          g("ux"), float(g("dt")))
        assert bit_equal(out, g("out")), i


@pytest.mark.parametrize("name", ["conv_c0_plane.npz", "conv_c1_plane.npz"])
def test_convection_plane_bit_exact(golden_dir, name):
    d = load(golden_dir, name)
    nP, nR = d["psd"].shape
    out = O.convection_2d(d["psd"], d["P"], d["R"], nP, nR, d["Pl"], d["Pu"], d["Rl"], d["Ru"],
                          tuple(int(t) for t in d["types"]), d["VP"], d["VR"], d["S"], d["L"], d["G"],
                          float(d["dt"]))
    assert bit_equal(out, d["out"])


def test_full_step_config0(golden_dir):
    from verb4d import workload as W
    d = load(golden_dir, "step_c0_rich.npz")
    w = W.config0(seed=int(d["seed"]), rich=bool(d["rich"]))
    out = O.step(w, w.psd0, nthreads=4)
    assert rel_err(out, d["out"]) < 1e-12


def test_oracle_matches_live_reference_when_present():
    """Where the reference harness is built (this container), run both on a
    fresh seed to guard against fixtures that happen to pass."""
    if not O.ref_available():
        pytest.skip("reference harness not built here")
    from verb4d import workload as W
    w = W.config0(seed=77, rich=True)
    a = O.step(w, w.psd0, 4, 1, 1, 0)
    b = O.step(w, w.psd0, 1, 1, 1, 0, lib="ref")
    assert bit_equal(a, b)
    a = O.step(w, w.psd0, 4, 0, 0, 1)
    b = O.step(w, w.psd0, 1, 0, 0, 1, lib="ref")
    assert node_rel_err(a, b) < 1e-12


@pytest.mark.parametrize("name", ["conv_lost_c0_plane.npz", "conv_lost_c1_plane.npz", "conv_fast_c0_plane.npz",
                                  "conv_fast_c1_plane.npz"])
def test_convection_build_options_bit_exact(golden_dir, name):
    """The oracle's Convection_2D with the reference's SAVE_PSD_LOST_CONV
    accumulator or FAST_CONVECTION rule against the reference built with
    that option."""
    d = load(golden_dir, name)
    nP, nR = d["psd"].shape
    fast = bool(int(d["fast"]))
    lost_in = d["lost_in"] if "lost_in" in d.files else None
    out, lost = O.convection_2d_ex(d["psd"], lost_in, d["P"], d["R"], nP, nR, d["Pl"], d["Pu"], d["Rl"], d["Ru"],
                                   tuple(int(t) for t in d["types"]), d["VP"], d["VR"], d["S"], d["L"], d["G"],
                                   float(d["dt"]), fast=fast)
    assert bit_equal(out, d["out"])
    if lost_in is not None:
        assert bit_equal(lost, d["lost_out"])


def test_build_options_change_the_result(golden_dir):
    """The option fixtures are not the plain build's answer: FAST_CONVECTION
    takes fewer outer sub-steps on the production plane at Kp 6, and the
    lost-PSD accumulator grows where the losses act."""
    d = load(golden_dir, "conv_fast_c1_plane.npz")
    nP, nR = d["psd"].shape
    plain = O.convection_2d(d["psd"], d["P"], d["R"], nP, nR, d["Pl"], d["Pu"], d["Rl"], d["Ru"],
                            tuple(int(t) for t in d["types"]), d["VP"], d["VR"], d["S"], d["L"], d["G"],
                            float(d["dt"]))
    assert not bit_equal(plain, d["out"])
    d = load(golden_dir, "conv_lost_c0_plane.npz")
    assert np.all(d["lost_out"] >= d["lost_in"]) and np.any(d["lost_out"] > d["lost_in"])
    # the plain PSD is the lost build's PSD
    nP, nR = d["psd"].shape
    plain = O.convection_2d(d["psd"], d["P"], d["R"], nP, nR, d["Pl"], d["Pu"], d["Rl"], d["Ru"],
                            tuple(int(t) for t in d["types"]), d["VP"], d["VR"], d["S"], d["L"], d["G"],
                            float(d["dt"]))
    assert bit_equal(plain, d["out"])


def test_lu_caching_build_matches_oracle_cache(golden_dir):
    """The reference built with LU_CACHING: its cached factors and pivots
    against the oracle's dgbtf2 cache, and the reuse call's result."""
    d = load(golden_dir, "local_lucache_node.npz")
    nx, ny = d["psd"].shape
    n = nx * ny
    kl = nx + 1
    ldab = 3 * kl + 1
    h = O.oracle_lib()
    args = [O._c(d[k]) for k in ("x", "y", "xl", "xu", "yl", "yu", "Dxx", "Dyy", "Dxy", "Dxy", "G", "S", "L")]
    cache = np.zeros(ldab * n)
    ipiv = np.zeros(n, dtype=np.int64)
    f = O._c(d["psd"]).copy()
    outs = []
    for recompute in (1, 0):
        st = h.orc_diffusion_2d(O._p(f), O._p(args[0]), O._p(args[1]), nx, ny, *[O._p(a) for a in args[2:6]],
                                *[int(v) for v in d["types"]], *[O._p(a) for a in args[6:]], float(d["dt"]),
                                float(d["sub_dt"]), O._p(cache), ipiv.ctypes.data_as(O._i64p), recompute)
        assert st == 0
        outs.append(f.copy())
        if recompute:
            assert np.array_equal(ipiv, d["ipiv"])
            assert rel_err(cache, d["cache"]) < 1e-12
    assert rel_err(outs[0], d["out1"]) < 1e-12
    assert rel_err(outs[1], d["out2"]) < 1e-12
