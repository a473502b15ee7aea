"""The reference's on-disk formats for numpy arrays, over the library's
host-side readers and writers (include/verb4d_hip.h, verb_plt_*,
verb_pltb_*, verb_grid_read, verb_tab_read, verb_queue_read,
verb_psd_lst_append), and an input folder as the reference's
ReadInitialData reads it (src/ReadInitialData.cpp:253-600): the grid, L*,
the coefficients, Jacobians, sources, losses and boundaries, each from
NAME.tab (a data file or .lst update list, scale, limits) or else from
NAME.plt / NAME.pltb.
"""
import ctypes
import os

import numpy as np

from . import TabLine, VerbError, lib

ASCII, BINARY = "ascii", "binary"


def _b(s):
    return os.fsencode(s) if s is not None else None


def _check(rc, what):
    if rc != 0:
        raise VerbError(rc, what)


def _sizes(shape):
    a = (ctypes.c_int * 4)(*([int(v) for v in shape] + [0] * (4 - len(shape))))
    return a


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def write_plt(path, a, name="", info=""):
    """Matrix4D/3D::writeToFile(filename, info) (src/Matrix.cpp:4556-4567, 3011-3024)."""
    a = _f64(a)
    sz = _sizes(a.shape)
    _check(lib().verb_plt_write(_b(path), a.ndim, ctypes.addressof(sz), a.ctypes.data, _b(name), _b(info)),
           "writing %s" % path)


def write_plt_grid(path, a, grids, grid_names=None, name=""):
    """writeToFile(filename, grid_w, grid_x, grid_y, grid_z) (:4833-4853, 3031-3050)."""
    a = _f64(a)
    gs = [_f64(g) for g in grids]
    assert len(gs) == a.ndim and all(g.shape == a.shape for g in gs)
    gp = (ctypes.c_void_p * 4)(*([g.ctypes.data for g in gs] + [None] * (4 - len(gs))))
    names = None
    if grid_names:
        names = (ctypes.c_char_p * 4)(*([_b(n) for n in grid_names] + [None] * (4 - len(grid_names))))
    sz = _sizes(a.shape)
    _check(lib().verb_plt_write_grid(_b(path), a.ndim, ctypes.addressof(sz), gp, names, a.ctypes.data, _b(name)),
           "writing %s" % path)


def plt_sizes(path):
    """The array shape a .plt file's ZONE line gives ([L,] K, J, I)."""
    s = (ctypes.c_int * 4)()
    nd = lib().verb_plt_sizes(_b(path), ctypes.addressof(s))
    if nd < 0:
        raise VerbError(nd, "reading the sizes of %s" % path)
    return tuple(s[:nd])


def read_plt(path, shape=None, column=1):
    """readFromFile(filename, column) (src/Matrix.cpp:4906-4965, 3059-3096);
    shape from the file's ZONE line when not given."""
    shape = tuple(shape) if shape is not None else plt_sizes(path)
    out = np.empty(shape, dtype=np.float64)
    sz = _sizes(shape)
    _check(lib().verb_plt_read(_b(path), int(column), len(shape), ctypes.addressof(sz), out.ctypes.data),
           "reading %s" % path)
    return out


def read_plt_checked(path, grids):
    """readFromFile(filename, grids...) (src/Matrix.cpp:5344-5413,
    3431-3477): every point's coordinates must agree with the grids within
    1e-6 in log10 (VerbError with .bad_index otherwise)."""
    gs = [_f64(g) for g in grids]
    shape = gs[0].shape
    out = np.empty(shape, dtype=np.float64)
    gp = (ctypes.c_void_p * 4)(*([g.ctypes.data for g in gs] + [None] * (4 - len(gs))))
    bad = ctypes.c_longlong(-1)
    sz = _sizes(shape)
    rc = lib().verb_plt_read_checked(_b(path), len(shape), ctypes.addressof(sz), gp, out.ctypes.data,
                                     ctypes.byref(bad))
    if rc != 0:
        e = VerbError(rc, "reading %s%s" % (path, "" if bad.value < 0 else
                                            ": grid mismatch at %s" % (np.unravel_index(bad.value, shape),)))
        e.bad_index = bad.value
        raise e
    return out


def write_pltb(path, a):
    """writeToBinaryFile (src/Matrix.cpp:5704-5726, 3781-3808)."""
    a = _f64(a)
    sz = _sizes(a.shape)
    _check(lib().verb_pltb_write(_b(path), a.ndim, ctypes.addressof(sz), a.ctypes.data), "writing %s" % path)


def pltb_sizes(path, ndim=4):
    with open(path, "rb") as f:
        return tuple(int(v) for v in np.frombuffer(f.read(4 * ndim), dtype=np.int32))


def read_pltb(path, shape=None, ndim=4):
    """readFromBinaryFile (src/Matrix.cpp:5792-5818, 3812-3838): the shape
    of the array it fills (the file's own sizes when not given)."""
    shape = tuple(shape) if shape is not None else pltb_sizes(path, ndim)
    out = np.empty(shape, dtype=np.float64)
    fs = (ctypes.c_int * 4)()
    sz = _sizes(shape)
    _check(lib().verb_pltb_read(_b(path), len(shape), ctypes.addressof(sz), out.ctypes.data, ctypes.addressof(fs)),
           "reading %s" % path)
    return out


def read_any(path, grids):
    """MatrixReadFromFile (src/UpdatableMatrix.cpp:97-124): .plt with the
    grid check, .pltb binary, anything else rejected."""
    if path.endswith(".plt"):
        return read_plt_checked(path, grids)
    if path.endswith(".pltb"):
        return read_pltb(path, shape=np.shape(grids[0]), ndim=len(grids))
    raise VerbError(-10, "Unknown data format: %s" % path)


def read_grid(folder):
    """(P, R, V, K) grids [P][R][V][K] of an input folder: grid.pltb, else
    grid.plt (src/ReadInitialData.cpp:253-346)."""
    folder = os.path.join(folder, "")
    s = (ctypes.c_int * 4)()
    _check(lib().verb_grid_read(_b(folder), ctypes.addressof(s), None, None, None, None, 0), "reading the grid")
    shape = tuple(s)
    g = [np.empty(shape, dtype=np.float64) for _ in range(4)]
    _check(lib().verb_grid_read(_b(folder), ctypes.addressof(s), *[a.ctypes.data for a in g], g[0].size),
           "reading the grid")
    return tuple(g)


def write_grid_pltb(folder, P, R, V, K):
    """grid.pltb as the reference reads it: the 4 sizes as int32, then the
    P, R, V and K grids."""
    P = _f64(P)
    with open(os.path.join(folder, "grid.pltb"), "wb") as f:
        f.write(np.asarray(P.shape, dtype=np.int32).tobytes())
        for g in (P, R, V, K):
            f.write(_f64(g).tobytes())


def append_psd_lst(output_folder, filename, binary, info):
    """Matrix4D::writeToLstFile (src/Matrix.cpp:5764-5785)."""
    _check(lib().verb_psd_lst_append(_b(os.path.join(output_folder, "") if output_folder else output_folder),
                                     _b(filename), 1 if binary else 0, _b(info)), "appending PSD.lst")


def read_tab(path):
    """The first line of a .tab file, split as UpdatableMatrix::readFromString
    splits it; None when the file does not exist."""
    t = TabLine()
    rc = lib().verb_tab_read(_b(path), ctypes.byref(t))
    if rc == -9:
        return None
    _check(rc, "reading %s" % path)
    dec = lambda v: v.decode()  # noqa: E731
    return {"has_line": bool(t.has_line), "data_kind": t.data_kind, "data": dec(t.data),
            "scale_present": bool(t.scale_present), "scale": dec(t.scale), "nlimits": t.nlimits,
            "from": [dec(t.from_[a].value) for a in range(4)], "to": [dec(t.to[a].value) for a in range(4)]}


def is_number(s):
    return bool(lib().verb_is_number(_b(s)))


def parse_number(s):
    return float(lib().verb_parse_number(_b(s)))


def read_queue(path, names=False):
    """[(time, value)] of a "time value" queue file (values are file names
    with names=True); [] when the file does not open, as the reference's
    queue stays empty."""
    n = ctypes.c_size_t()
    rc = lib().verb_queue_read(_b(path), None, None, None, 0, 0, ctypes.byref(n))
    if rc == -9:
        return []
    _check(rc, "reading %s" % path)
    k = n.value
    times = np.empty(max(k, 1), dtype=np.float64)
    if names:
        L = 4096
        buf = ctypes.create_string_buffer(L * max(k, 1))
        _check(lib().verb_queue_read(_b(path), times.ctypes.data, None, buf, L, k, ctypes.byref(n)),
               "reading %s" % path)
        raw = buf.raw
        vals = [raw[i * L:(i + 1) * L].split(b"\0", 1)[0].decode() for i in range(k)]
    else:
        v = np.empty(max(k, 1), dtype=np.float64)
        _check(lib().verb_queue_read(_b(path), times.ctypes.data, v.ctypes.data, None, 0, k, ctypes.byref(n)),
               "reading %s" % path)
        vals = [float(x) for x in v[:k]]
    return [(float(times[i]), vals[i]) for i in range(k)]


def _resolve(p, cwd):
    return p if (os.path.isabs(p) or cwd is None) else os.path.join(cwd, p)


def field_spec(tab, grids, cwd=None):
    """The update description of one .tab line (read_tab's dict): the data
    array (or None), the .lst file queue, the scale (constant and queue) and
    the per-axis limits, in the form verb4d.timeloop.UpdatableField takes.
    Relative file names resolve against cwd (the reference opens them
    relative to its working directory)."""
    spec = {"original": None, "files": [], "scale": 1.0, "scales": [], "limit_init": {}, "limit_from": {},
            "limit_to": {}, "limited": False}
    if tab is None or not tab["has_line"]:
        return spec
    if tab["data_kind"] == 1:
        spec["original"] = read_any(_resolve(tab["data"], cwd), grids)
    elif tab["data_kind"] == 2:
        spec["files"] = [(t, v if v == "-" else _resolve(v, cwd))
                         for t, v in read_queue(_resolve(tab["data"], cwd), names=True)]
    sc = tab["scale"]
    if tab["scale_present"] and sc not in ("", "-"):
        # (an empty field, read past trailing blanks, scales nothing)
        spec["scale_active"] = True
        if is_number(sc):
            spec["scale"] = parse_number(sc)
        else:
            spec["scales"] = read_queue(_resolve(sc, cwd))
    for a in range(tab["nlimits"]):
        spec["limited"] = True
        lo, hi = -1e99, 1e99
        f, t = tab["from"][a], tab["to"][a]
        if f != "-":
            if is_number(f):
                lo = parse_number(f)
            else:
                spec["limit_from"][a] = read_queue(_resolve(f, cwd))
        if t != "-":
            if is_number(t):
                hi = parse_number(t)
            else:
                spec["limit_to"][a] = read_queue(_resolve(t, cwd))
        spec["limit_init"][a] = (lo, hi)
    return spec


# ReadInitialData's names (src/ReadInitialData.cpp:350-560): the field
# name of this library, the file base name, and the grids the reference
# checks the file against
FIELDS_4D = [("DLL", "DLL"), ("DVV", "DVV"), ("DKK", "DKK"), ("DVK", "DVK"), ("VP", "VP"), ("VR", "VR"),
             ("G_local", "G_local"), ("G_radial", "G_radial"), ("G_conv", "G_conv"), ("sources", "Sources"),
             ("losses_local", "Losses_local"), ("losses_radial", "Losses_radial"), ("losses_conv", "Losses_conv")]
BOUNDARIES = [("Rl", "Rl_BC", 1, 0), ("Ru", "Ru_BC", 1, -1), ("Ll", "Ll_BC", 1, 0), ("Lu", "Lu_BC", 1, -1),
              ("Vl", "Vl_BC", 2, 0), ("Vu", "Vu_BC", 2, -1), ("Kl", "Kl_BC", 3, 0), ("Ku", "Ku_BC", 3, -1)]


def read_field(folder, base, grids, io_method=ASCII, cwd=None, required=True):
    """One field as ReadInitialData reads it: BASE.tab if it exists (its
    update description), else BASE.plt (grid-checked) or BASE.pltb.
    Returns (array or None, spec or None)."""
    tab = read_tab(os.path.join(folder, base + ".tab"))
    if tab is not None:
        spec = field_spec(tab, grids, cwd)
        return spec["original"], spec
    path = os.path.join(folder, base + (".plt" if io_method == ASCII else ".pltb"))
    if not os.path.exists(path):
        if required:
            raise VerbError(-9, "%s not found" % path)
        return None, None
    if io_method == ASCII:
        return read_plt_checked(path, grids), None
    return read_pltb(path, shape=np.shape(grids[0]), ndim=len(grids)), None


class FileWorkload:
    """The arrays of an input folder in the layout verb4d.device.DeviceGrid
    takes (the attributes of verb4d.workload.Workload): grid, L*, initial
    PSD, coefficients, Jacobians, sources, losses and boundaries. Fields
    absent from the folder take the reference's defaults where it has one
    (G_* = 1, sources and losses 0). `specs` holds each .tab field's update
    description for verb4d.timeloop."""

    def __init__(self, folder, initial_psd="PSD0", dt=300.0, sub_dt=None, io_method=ASCII, types=None,
                 cwd=None, include_boundary=1, radial_sources=0.0, local_sources=1.0):
        P, R, V, K = read_grid(folder)
        self.nP, self.nR, self.nV, self.nK = P.shape
        self.shape = P.shape
        self.dt = float(dt)
        self.sub_dt = float(dt if sub_dt is None else sub_dt)
        g4 = (P, R, V, K)
        self.f = {"Pg": P, "Rg": R, "Vg": V, "Kg": K}
        self.specs = {}
        L, spec = read_field(folder, "Lstar", g4, io_method, cwd, required=False)
        self.f["Lg"] = L if L is not None else R
        if spec is not None:
            self.specs["Lstar"] = spec
        # the initial PSD: .pltb when it exists, else .plt (without the grid
        # check, src/ReadInitialData.cpp:357-362)
        b = os.path.join(folder, initial_psd)
        self.psd0 = read_pltb(b + ".pltb", shape=self.shape) if os.path.exists(b + ".pltb") else \
            read_plt(b + ".plt", shape=self.shape)
        defaults = {"G_local": 1.0, "G_radial": 1.0, "G_conv": 1.0, "sources": 0.0, "losses_local": 0.0,
                    "losses_radial": 0.0, "losses_conv": 0.0}
        for name, base in FIELDS_4D:
            grids = (P, self.f["Lg"], V, K) if name == "DLL" else g4
            if name == "losses_local":
                # Losses.tab first, for older inputs (:476-481)
                arr, spec = read_field(folder, "Losses", grids, io_method, cwd, required=False)
                if arr is None and spec is None:
                    arr, spec = read_field(folder, base, grids, io_method, cwd, required=False)
            else:
                arr, spec = read_field(folder, base, grids, io_method, cwd, required=name not in defaults)
            if spec is not None:
                self.specs[name] = spec
            if arr is None:
                # a .tab line without a data file leaves the reference's array
                # unset until its first .lst entry (new double[], never
                # filled); zeros here
                arr = np.zeros(self.shape) if spec is not None else np.full(self.shape, defaults[name])
            self.f[name] = np.ascontiguousarray(arr)
        self.bc = {"Pl": np.full((self.nR, self.nV, self.nK), -1e99), "Pu": np.full((self.nR, self.nV, self.nK), -1e99)}
        for name, base, axis, idx in BOUNDARIES:
            grids = tuple(np.ascontiguousarray(np.take(G, idx, axis=axis)) for i, G in enumerate(g4) if i != axis)
            arr, spec = read_field(folder, base, grids, io_method, cwd, required=False)
            if spec is not None:
                self.specs[name] = spec
            self.bc[name] = np.ascontiguousarray(arr) if arr is not None else np.zeros(grids[0].shape)
        self.types = dict(Pl=0, Pu=0, Rl=1, Ru=1, Ll=1, Lu=1, Vl=1, Vu=1, Kl=1, Ku=1)
        self.types.update(types or {})
        self.include_boundary = include_boundary
        self.radial_sources = radial_sources
        self.local_sources = local_sources
        self.kp = float("nan")

    @property
    def points(self):
        return self.nP * self.nR * self.nV * self.nK

    def fill_struct(self, g, ptr_of):
        g.nP, g.nR, g.nV, g.nK = self.nP, self.nR, self.nV, self.nK
        for k, v in self.f.items():
            setattr(g, k, ptr_of(k, v))
        for k, v in self.bc.items():
            setattr(g, k, ptr_of(k, v))
        for k, t in self.types.items():
            setattr(g, k + "_type", t)
        g.dt, g.sub_dt = self.dt, self.sub_dt
        g.radial_sources, g.local_sources = self.radial_sources, self.local_sources
        g.include_boundary = self.include_boundary
        return g


def write_inputs(folder, w, io_method=ASCII, initial_psd="PSD0", grid_binary=True):
    """An input folder for a workload's arrays (grid, L*, initial PSD,
    coefficients, boundaries), the layout FileWorkload reads."""
    os.makedirs(folder, exist_ok=True)
    P, R, V, K = (w.f[k] for k in ("Pg", "Rg", "Vg", "Kg"))
    if grid_binary:
        write_grid_pltb(folder, P, R, V, K)
    else:
        # grid.plt: the four grids as the first four columns
        write_plt_grid(os.path.join(folder, "grid.plt"), w.f["Vg"], (P, R, V, K), name="unused")
    ext = ".plt" if io_method == ASCII else ".pltb"

    def put(base, a, grids):
        p = os.path.join(folder, base + ext)
        if io_method == ASCII:
            write_plt_grid(p, a, grids)
        else:
            write_pltb(p, a)

    g4 = (P, R, V, K)
    put("Lstar", w.f["Lg"], g4)
    if io_method == ASCII:
        write_plt(os.path.join(folder, initial_psd + ".plt"), w.psd0)
    else:
        write_pltb(os.path.join(folder, initial_psd + ".pltb"), w.psd0)
    for name, base in FIELDS_4D:
        put(base, w.f[name], (P, w.f["Lg"], V, K) if name == "DLL" else g4)
    for name, base, axis, idx in BOUNDARIES:
        grids = tuple(np.ascontiguousarray(np.take(G, idx, axis=axis)) for i, G in enumerate(g4) if i != axis)
        put(base, w.bc[name], grids)


__all__ = ["write_plt", "write_plt_grid", "plt_sizes", "read_plt", "read_plt_checked", "write_pltb", "read_pltb",
           "read_any", "read_grid", "write_grid_pltb", "append_psd_lst", "read_tab", "is_number", "parse_number",
           "read_queue", "field_spec", "read_field", "FileWorkload", "write_inputs", "ASCII", "BINARY"]
