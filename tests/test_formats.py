"""The on-disk formats (verb-4d_amd/csrc/formats.cpp through verb4d.formats)
against the reference's own Matrix3D/Matrix4D writers and readers
(oracle/_ref/libverbref.so: src/Matrix.cpp compiled from the reference
tree) where those run on the input, and against what the formats fix
otherwise (the grid.pltb layout of src/ReadInitialData.cpp:261-290, the
.tab splitting rules of src/UpdatableMatrix.cpp:392-551, is_number and the
istream number prefix).

Bars: files byte-identical; arrays read bit-identical; the same grid
check. Host code only: no device work.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
import verb4d
from verb4d import formats as F
from verb4d import workload as W

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference harness not built")


def _ip(shape):
    return (ctypes.c_int * 4)(*(list(shape) + [0] * (4 - len(shape))))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def sample(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) * 10.0 ** rng.uniform(-30, 30, shape)
    flat = a.reshape(-1)
    flat[:6] = [0.0, -0.0, 1.0, -1.0, 5e-324, 1.7976931348623157e308]
    return a


def grids4(shape):
    P = np.broadcast_to(np.linspace(0.0, 6.283185307179586, shape[0])[:, None, None, None], shape) + 0.1
    R = np.broadcast_to(np.linspace(1.5, 6.6, shape[1])[None, :, None, None], shape)
    V = np.broadcast_to(np.logspace(-2, 1, shape[2])[None, None, :, None], shape)
    K = np.broadcast_to(np.linspace(0.1, 1.57, shape[3])[None, None, None, :], shape)
    return tuple(np.ascontiguousarray(g) for g in (P, R, V, K))


@needs_ref
@pytest.mark.parametrize("name,info", [("", ""), ("PSD", "t=1.5 days")])
def test_plt4_write_matches_reference_bytes(tmp_path, name, info):
    shape = (3, 4, 5, 6)
    a = sample(shape, 1)
    ours, ref = str(tmp_path / "a.plt"), str(tmp_path / "b.plt")
    F.write_plt(ours, a, name=name, info=info)
    O.ref_lib().ref_plt4_write(ref.encode(), _ip(shape), _dp(a), name.encode(), info.encode())
    # the title defaults to the file name: compare with the names swapped in
    got = open(ours, "rb").read().replace(ours.encode(), b"@")
    want = open(ref, "rb").read().replace(ref.encode(), b"@")
    assert got == want


@needs_ref
def test_plt4_grid_write_and_reads_match_reference(tmp_path):
    shape = (3, 4, 5, 6)
    a = sample(shape, 2)
    g = grids4(shape)
    ours, ref = str(tmp_path / "a.plt"), str(tmp_path / "b.plt")
    F.write_plt_grid(ours, a, g, name="DVV")
    O.ref_lib().ref_plt4_write_grid(ref.encode(), _ip(shape), *[_dp(x) for x in g], _dp(a), b"DVV")
    got = open(ours, "rb").read().replace(ours.encode(), b"@")
    want = open(ref, "rb").read().replace(ref.encode(), b"@")
    assert got == want
    # both readers on the reference's file: the checked read and column 5
    want = np.empty(shape)
    O.ref_lib().ref_plt4_read_checked(ref.encode(), _ip(shape), *[_dp(x) for x in g], _dp(want))
    assert np.array_equal(bits(F.read_plt_checked(ref, g)), bits(want))
    O.ref_lib().ref_plt4_read_column(ref.encode(), 5, _ip(shape), _dp(want))
    assert np.array_equal(bits(F.read_plt(ref, shape, column=5)), bits(want))
    for c in range(1, 5):
        O.ref_lib().ref_plt4_read_column(ref.encode(), c, _ip(shape), _dp(want))
        assert np.array_equal(bits(F.read_plt(ref, shape, column=c)), bits(want))
    assert F.plt_sizes(ref) == shape


@needs_ref
def test_plt3_and_pltb_match_reference(tmp_path):
    s3 = (4, 5, 6)
    a = sample(s3, 3)
    G = grids4((4, 1, 5, 6))
    g3 = tuple(np.ascontiguousarray(G[i][:, 0]) for i in (0, 2, 3))  # a boundary's (P, V, K) grids
    assert all(x.shape == s3 for x in g3)
    for kind in ("plain", "grid"):
        ours, ref = str(tmp_path / "a3.plt"), str(tmp_path / "b3.plt")
        if kind == "plain":
            F.write_plt(ours, a, name="Rl", info="")
            O.ref_lib().ref_plt3_write(ref.encode(), _ip(s3), _dp(a), b"Rl", b"")
        else:
            F.write_plt_grid(ours, a, g3)
            O.ref_lib().ref_plt3_write_grid(ref.encode(), _ip(s3), *[_dp(x) for x in g3], _dp(a), b"")
        assert open(ours, "rb").read().replace(ours.encode(), b"@") == open(ref, "rb").read().replace(
            ref.encode(), b"@")
    want = np.empty(s3)
    O.ref_lib().ref_plt3_read_checked(ref.encode(), _ip(s3), *[_dp(x) for x in g3], _dp(want))
    assert np.array_equal(bits(F.read_plt_checked(ref, g3)), bits(want))
    # binary, 3-D and 4-D
    for shape, w, r in (((4, 5, 6), "ref_pltb3_write", "ref_pltb3_read"),
                        ((3, 4, 5, 6), "ref_pltb4_write", "ref_pltb4_read")):
        b = sample(shape, 4)
        ours, ref = str(tmp_path / "a.pltb"), str(tmp_path / "b.pltb")
        F.write_pltb(ours, b)
        getattr(O.ref_lib(), w)(ref.encode(), _ip(shape), _dp(b))
        assert open(ours, "rb").read() == open(ref, "rb").read()
        want = np.empty(shape)
        getattr(O.ref_lib(), r)(ours.encode(), _ip(shape), _dp(want))
        got = F.read_pltb(ref, ndim=len(shape))
        assert got.shape == shape
        assert np.array_equal(bits(got), bits(want)) and np.array_equal(bits(got), bits(b))


@needs_ref
def test_psd_lst_matches_reference(tmp_path):
    out = str(tmp_path) + "/"
    for binary, info in ((0, "0.5"), (1, "1.25"), (0, "")):
        F.append_psd_lst(str(tmp_path), out + "PSD_%s" % info, binary, info)
    ours = open(out + "PSD.lst").read()
    os.remove(out + "PSD.lst")
    for binary, info in ((0, "0.5"), (1, "1.25"), (0, "")):
        O.ref_lib().ref_lst_append((out + "PSD_%s" % info).encode(), binary, info.encode(), out.encode())
    assert ours == open(out + "PSD.lst").read()
    assert ours.splitlines()[1] == "1.25\tPSD_1.25.pltb"


def test_grid_check_tolerance_and_errors(tmp_path):
    shape = (2, 3, 4, 5)
    a = sample(shape, 5)
    g = grids4(shape)
    p = str(tmp_path / "g.plt")
    F.write_plt_grid(p, a, g)
    # 1e-6 decades: 1 + 1e-6 relative passes, 1 + 1e-5 does not
    gs = [x.copy() for x in g]
    gs[2].reshape(-1)[37] *= 1.000001
    F.read_plt_checked(p, gs)
    gs[2].reshape(-1)[37] *= 1.00001
    with pytest.raises(verb4d.VerbError) as e:
        F.read_plt_checked(p, gs)
    assert e.value.status == -10 and e.value.bad_index == 37
    # more data than the sizes ("Check grid size") and a short file
    with pytest.raises(verb4d.VerbError):
        F.read_plt(p, (2, 3, 4, 4), column=5)
    with pytest.raises(verb4d.VerbError):
        F.read_plt(p, (2, 3, 4, 6), column=5)
    # a value the stream cannot read
    txt = open(p).read().splitlines()
    txt[5] = txt[5].rsplit("\t", 1)[0] + "\tnan"
    open(p, "w").write("\n".join(txt) + "\n")
    with pytest.raises(verb4d.VerbError):
        F.read_plt(p, shape, column=5)
    with pytest.raises(verb4d.VerbError) as e:
        F.read_plt(str(tmp_path / "missing.plt"), shape)
    assert e.value.status == -9


def test_grid_files(tmp_path):
    """grid.pltb: 4 int32 sizes then the P, R, V, K grids
    (src/ReadInitialData.cpp:261-290); grid.plt: the ZONE sizes (I=K, J=V,
    K=R, L=P) and the grids as columns 1-4 (:292-343)."""
    shape = (3, 4, 5, 6)
    g = grids4(shape)
    d = tmp_path / "b"
    d.mkdir()
    with open(d / "grid.pltb", "wb") as f:
        f.write(np.array(shape, dtype=np.int32).tobytes() + b"".join(x.tobytes() for x in g))
    got = F.read_grid(str(d))
    assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(got, g))
    a = tmp_path / "a"
    a.mkdir()
    F.write_plt_grid(str(a / "grid.plt"), np.zeros(shape), g, grid_names=["P", "R", "V", "K"])
    got = F.read_grid(str(a))
    for x, y in zip(got, g):
        assert np.max(np.abs(x - y) / np.abs(y)) < 5e-7  # "%e" keeps 7 digits
        assert x.shape == shape


def test_tab_lines(tmp_path):
    def tab(text):
        p = tmp_path / "x.tab"
        p.write_text(text)
        return F.read_tab(str(p))

    t = tab("DVV.plt 2.0 0 12 - 5\n")
    assert (t["data_kind"], t["data"], t["scale"], t["nlimits"]) == (1, "DVV.plt", "2.0", 2)
    assert (t["from"][:2], t["to"][:2]) == (["0", "-"], ["12", "5"])
    t = tab("updates.lst")
    assert t["data_kind"] == 2 and not t["scale_present"] and t["nlimits"] == 0
    # a field after trailing blanks is read, and empty (the stream has not ended)
    t = tab("DVV.plt 3 \n")
    assert t["scale"] == "3" and t["nlimits"] == 1 and t["from"][0] == "" and t["to"][0] == ""
    t = tab("- scale.txt")
    assert t["data_kind"] == 0 and t["scale"] == "scale.txt"
    # the reference's suffix test: names shorter than ".lst" count as lists
    assert tab("abc")["data_kind"] == 2
    assert not tab("# DVV.plt 2")["has_line"] and not tab("\nDVV.plt")["has_line"]
    assert F.read_tab(str(tmp_path / "none.tab")) is None
    assert [F.is_number(s) for s in ("1e5", "-1.5e-3", "1E5", "", "12x", "+.5")] == [1, 1, 0, 0, 0, 1]
    assert [F.parse_number(s) for s in ("1.5.3", "--1", "2e3", "-.25", "1e")] == [1.5, 0.0, 2000.0, -0.25, 0.0]


def test_queues_and_field_spec(tmp_path):
    (tmp_path / "scale.txt").write_text("0 1.5\n0.25 2.0\n\n1.0 1e-1\n")
    (tmp_path / "upd.lst").write_text("0.5 a.pltb\n1.5 -\n")
    assert F.read_queue(str(tmp_path / "scale.txt")) == [(0.0, 1.5), (0.25, 2.0), (1.0, 0.1)]
    assert F.read_queue(str(tmp_path / "upd.lst"), names=True) == [(0.5, "a.pltb"), (1.5, "-")]
    assert F.read_queue(str(tmp_path / "missing.txt")) == []
    shape = (2, 3, 4, 5)
    g = grids4(shape)
    F.write_pltb(str(tmp_path / "a.pltb"), np.ones(shape))
    (tmp_path / "f.tab").write_text("upd.lst scale.txt - - 2 3\n")
    spec = F.field_spec(F.read_tab(str(tmp_path / "f.tab")), g, cwd=str(tmp_path))
    assert spec["files"] == [(0.5, str(tmp_path / "a.pltb")), (1.5, "-")]
    assert spec["scales"] == [(0.0, 1.5), (0.25, 2.0), (1.0, 0.1)]
    assert spec["limit_init"] == {0: (-1e99, 1e99), 1: (2.0, 3.0)} and spec["limited"]
    # driven through the oracle's UpdatableMatrix::update restatement
    import timeloop_oracle as OT
    u = OT.UpdatableArray(np.zeros(shape), files=spec["files"], scales=spec["scales"], scale=spec["scale"],
                          limit_from=spec["limit_from"], limit_to=spec["limit_to"], limit_init=spec["limit_init"],
                          grids=g, loader=lambda p: F.read_any(p, g))
    assert u.update(0.0) is True and np.all(u.current == 0.0)
    assert u.update(0.6) is True
    inside = (g[1] >= 2.0) & (g[1] <= 3.0)
    assert np.array_equal(u.current, np.where(inside, 2.0, 0.0))
    assert u.update(1.6) is True and np.array_equal(u.current, np.where(inside, 0.1, 0.0))


@pytest.mark.parametrize("io", [F.BINARY, F.ASCII])
def test_input_folder_round_trip(tmp_path, io):
    w = W.Workload(nP=4, nR=6, nV=7, nK=8, seed=9, rich=True)
    w.f["Lg"] = W.lstar(w, 0.05)
    F.write_inputs(str(tmp_path), w, io_method=io)
    fw = F.FileWorkload(str(tmp_path), dt=w.dt, io_method=io, types=w.types)
    names = ["Pg", "Rg", "Vg", "Kg", "Lg", "DLL", "DVV", "DKK", "DVK", "VP", "VR", "G_local", "G_radial", "G_conv",
             "sources", "losses_local", "losses_radial", "losses_conv"]
    for n in names:
        a, b = fw.f[n], w.f[n]
        if io == F.BINARY or n in ("Pg", "Rg", "Vg", "Kg"):
            assert np.array_equal(bits(a), bits(np.ascontiguousarray(b))), n
        else:
            assert np.all(np.abs(a - b) <= 5e-7 * np.abs(b)), n
    for n in ("Rl", "Ru", "Vl", "Vu", "Kl", "Ku"):
        assert fw.bc[n].shape == w.bc[n].shape
        assert np.all(np.abs(fw.bc[n] - w.bc[n]) <= 5e-7 * np.abs(w.bc[n])), n
    assert np.all(np.abs(fw.psd0 - w.psd0) <= 5e-7 * np.abs(w.psd0))
