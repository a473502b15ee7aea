"""verb_diffusion_2d_noalloc (static per-GPU scratch, no device allocation)
against verb_diffusion_2d, through the C-ABI: the same kernels on the same
inputs must give bit-identical PSD (and LU cache), on two calls in a row
whose second is smaller, plain and LU-caching variants; and a node larger
than the static scratch is rejected on the host."""
import os

import numpy as np
import pytest

import verb4d
from parity import bit_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return verb4d.Context()


def _args(d):
    nx, ny = d["psd"].shape
    t = [int(v) for v in d["types"]]
    return (d["x"], d["y"], nx, ny, d["xl"], d["xu"], d["yl"], d["yu"], *t, d["Dxx"], d["Dyy"], d["Dxy"], d["Dxy"],
            d["G"], d["S"], d["L"], float(d["dt"]), float(d["sub_dt"]))


def test_noalloc_matches_pooled_entry_two_calls(ctx, golden_dir):
    for name in ("local_c1_node.npz", "local_c0_rich_substeps.npz"):  # second call smaller
        print("noalloc", name, flush=True)
        d = np.load(os.path.join(golden_dir, name))
        ref = verb4d.Diffusion_2D(d["psd"], *_args(d), ctx=ctx)
        got = verb4d.Diffusion_2D_noalloc(d["psd"], *_args(d), ctx=ctx)
        assert bit_equal(got, ref), name


def test_noalloc_lu_caching_matches_pooled_entry(ctx, golden_dir):
    for name in ("local_c0_node.npz", "local_lucache_node.npz"):  # second call smaller
        print("noalloc lucache", name, flush=True)
        d = np.load(os.path.join(golden_dir, name))
        nx, ny = d["psd"].shape
        n = nx * ny
        ldab = 3 * (nx + 1) + 1
        c1, p1 = np.zeros(ldab * n), np.zeros(n, dtype=np.int64)
        c2, p2 = np.zeros(ldab * n), np.zeros(n, dtype=np.int64)
        r1 = verb4d.Diffusion_2D(d["psd"], *_args(d), lu_cache=c1, index_cache=p1, recompute_lu=True, ctx=ctx)
        g1 = verb4d.Diffusion_2D_noalloc(d["psd"], *_args(d), lu_cache=c2, index_cache=p2, recompute_lu=True,
                                         ctx=ctx)
        assert bit_equal(g1, r1) and bit_equal(c2, c1) and np.array_equal(p2, p1), name
        r2 = verb4d.Diffusion_2D(r1, *_args(d), lu_cache=c1, index_cache=p1, recompute_lu=False, ctx=ctx)
        g2 = verb4d.Diffusion_2D_noalloc(g1, *_args(d), lu_cache=c2, index_cache=p2, recompute_lu=False, ctx=ctx)
        assert bit_equal(g2, r2), name


def test_noalloc_rejects_node_beyond_static_scratch(ctx):
    nx, ny = 151, 121  # ~68 MB of band factors: more than the 32 MiB array
    z2 = np.zeros((nx, ny))
    with pytest.raises(verb4d.VerbError) as e:
        verb4d.Diffusion_2D_noalloc(z2, z2, z2, nx, ny, np.zeros(ny), np.zeros(ny), np.zeros(nx), np.zeros(nx), 1, 1,
                                    1, 1, z2, z2, z2, z2, z2, z2, z2, 300.0, 300.0, ctx=ctx)
    assert e.value.status == -1
