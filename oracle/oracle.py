"""ctypes access to the CPU oracle (liboracle.so) and to the reference
harness (_ref/libverbref.so, built from the PRBEM/VERB-4D sources).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / the baseline. The product
(verb-4d_amd/) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libverbref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_I = ctypes.c_int
_ip = ctypes.POINTER(ctypes.c_int)
_D = ctypes.c_double


class Grid(ctypes.Structure):
    """Same layout as the oracle's orc_grid and the harness's ref_grid."""

    _fields_ = (
        [("nP", _I), ("nR", _I), ("nV", _I), ("nK", _I)]
        + [(n, ctypes.c_void_p) for n in ("Pg", "Rg", "Vg", "Kg", "Lg")]
        + [(n, ctypes.c_void_p) for n in ("DLL", "DVV", "DKK", "DVK", "VP", "VR")]
        + [
            (n, ctypes.c_void_p)
            for n in ("G_local", "G_radial", "G_conv", "sources", "losses_local", "losses_radial", "losses_conv")
        ]
        + [(n, ctypes.c_void_p) for n in ("Pl", "Pu", "Rl", "Ru", "Ll", "Lu", "Vl", "Vu", "Kl", "Ku")]
        + [(n + "_type", _I) for n in ("Pl", "Pu", "Rl", "Ru", "Ll", "Lu", "Vl", "Vu", "Kl", "Ku")]
        + [("dt", _D), ("sub_dt", _D), ("radial_sources", _D), ("local_sources", _D), ("include_boundary", _I)]
    )


def build():
    """make -C oracle (the oracle always; the reference harness when the
    reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_oracle = None
_ref = {}
# reference harness builds: the reference alone, and the reference with the
# product's drop-ins linked in (built only where the product library exists)
REF_VARIANTS = {
    "ref": REF_LIB,
    "hipops": os.path.join(HERE, "_ref", "libverbref_hipops.so"),
    "hiplapack": os.path.join(HERE, "_ref", "libverbref_hiplapack.so"),
    # the reference's remapping loop calling the GPU Cubic1D drop-in
    "hipremap": os.path.join(HERE, "_ref", "libverbref_hipremap.so"),
    # the reference built with its compile-time options on this path
    "lucache": os.path.join(HERE, "_ref", "libverbref_lucache.so"),
    "lost": os.path.join(HERE, "_ref", "libverbref_lost.so"),
    "fast": os.path.join(HERE, "_ref", "libverbref_fast.so"),
}


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        h = ctypes.CDLL(ORACLE_LIB)
        h.orc_assemble_2d.argtypes = [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 7 + [
            _D, _dp, _dp, _dp]
        h.orc_pack_band.argtypes = [_I, _I, _dp, _dp]
        h.orc_dgbtf2.argtypes = [_I, _I, _I, _I, _dp, _I, _i64p]
        h.orc_dgbtrs.argtypes = [_I, _I, _I, _dp, _I, _i64p, _dp]
        h.orc_diffusion_2d.argtypes = [_dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 7 + [
            _D, _D, _dp, _i64p, _I]
        h.orc_tridag.argtypes = [_dp, _dp, _dp, _dp, _dp, _I]
        h.orc_diffusion_1d.argtypes = [_dp, _dp, _I, _D, _D, _I, _I, _dp, _dp, _dp, _dp, _D]
        h.orc_quickest6.argtypes = [_dp, _dp, _I, _D, _D, _I, _I, _dp, _D]
        h.orc_convection_2d.argtypes = [_dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 5 + [_D]
        h.orc_convection_2d_ex.argtypes = [_dp, _dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [
            _dp] * 5 + [_D, _I]
        h.orc_step.argtypes = [ctypes.POINTER(Grid), _dp, _I, _I, _I, _I]
        h.orc_cubic1d.argtypes = [_I, _dp, _dp, _dp]
        h.orc_remap.argtypes = [_dp, _dp, _dp, _I, _I, _I, _I]
        h.orc_grid_struct_size.restype = ctypes.c_size_t
        assert h.orc_grid_struct_size() == ctypes.sizeof(Grid)
        _oracle = h
    return _oracle


def ref_available(variant="ref"):
    return os.path.exists(REF_VARIANTS[variant])


def ref_lib(variant="ref"):
    if variant not in _ref:
        path = REF_VARIANTS[variant]
        if not os.path.exists(path):
            raise RuntimeError("reference harness not built (%s)" % path)
        h = ctypes.CDLL(path)
        # each build exports the entries its options allow
        sigs = {
            "ref_diffusion_2d": [_dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 7 + [_D, _D],
            "ref_diffusion_2d_cached": [_dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 7 + [
                _D, _D, _dp, _i64p, _I],
            "ref_assemble_2d": [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 7 + [
                _D, _dp, _dp, _dp],
            "ref_diffusion_1d": [_dp, _dp, _I, _D, _D, _I, _I, _dp, _dp, _dp, _dp, _D],
            "ref_tridag": [_dp, _dp, _dp, _dp, _dp, _I],
            "ref_quickest6": [_dp, _dp, _I, _D, _D, _I, _I, _dp, _D],
            "ref_convection_2d": [_dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [_dp] * 5 + [_D],
            "ref_convection_2d_lost": [_dp, _dp, _dp, _dp, _I, _I, _dp, _dp, _dp, _dp, _I, _I, _I, _I] + [
                _dp] * 5 + [_D],
            "ref_step_sample": [ctypes.POINTER(Grid), _dp, _I, _I, _I, _I, ctypes.c_long, ctypes.c_long,
                                ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long],
            "ref_step_local_cached": [ctypes.POINTER(Grid), _dp, _I, _dp, _i64p, _I, ctypes.c_long,
                                      ctypes.c_long],
            "ref_cubic1d": [_I, _dp, _dp, _dp],
            "ref_plt4_write": [ctypes.c_char_p, _ip, _dp, ctypes.c_char_p, ctypes.c_char_p],
            "ref_plt4_write_grid": [ctypes.c_char_p, _ip, _dp, _dp, _dp, _dp, _dp, ctypes.c_char_p],
            "ref_plt4_read_column": [ctypes.c_char_p, _I, _ip, _dp],
            "ref_plt4_read_checked": [ctypes.c_char_p, _ip, _dp, _dp, _dp, _dp, _dp],
            "ref_pltb4_write": [ctypes.c_char_p, _ip, _dp],
            "ref_pltb4_read": [ctypes.c_char_p, _ip, _dp],
            "ref_lst_append": [ctypes.c_char_p, _I, ctypes.c_char_p, ctypes.c_char_p],
            "ref_plt3_write": [ctypes.c_char_p, _ip, _dp, ctypes.c_char_p, ctypes.c_char_p],
            "ref_plt3_write_grid": [ctypes.c_char_p, _ip, _dp, _dp, _dp, _dp, ctypes.c_char_p],
            "ref_plt3_read_checked": [ctypes.c_char_p, _ip, _dp, _dp, _dp, _dp],
            "ref_pltb3_write": [ctypes.c_char_p, _ip, _dp],
            "ref_pltb3_read": [ctypes.c_char_p, _ip, _dp],
            "ref_remap": [_dp, _dp, _dp, _I, _I, _I, _I],
        }
        for name, args in sigs.items():
            if hasattr(h, name):
                getattr(h, name).argtypes = args
        if hasattr(h, "ref_grid_struct_size"):
            h.ref_grid_struct_size.restype = ctypes.c_size_t
            assert h.ref_grid_struct_size() == ctypes.sizeof(Grid)
        _ref[variant] = h
    return _ref[variant]


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_dp)


# --------------------------------------------------------------------------
# Oracle calls (numpy in, numpy out; input arrays are not modified)
# --------------------------------------------------------------------------
def _local_args(x, y, nx, ny, xl, xu, yl, yu, Dxx, Dyy, Dxy, Dyx, G, S, L):
    arrs = [_c(v) for v in (x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, Dyx, G, S, L)]
    return arrs


def diffusion_2d(psd, x, y, nx, ny, xl, xu, yl, yu, types, Dxx, Dyy, Dxy, Dyx, G, S, L, dt, sub_dt, lib="oracle"):
    out = _c(psd).copy()
    a = _local_args(x, y, nx, ny, xl, xu, yl, yu, Dxx, Dyy, Dxy, Dyx, G, S, L)
    if lib == "oracle":
        st = oracle_lib().orc_diffusion_2d(_p(out), _p(a[0]), _p(a[1]), nx, ny, _p(a[2]), _p(a[3]), _p(a[4]),
                                           _p(a[5]), *types, *[_p(v) for v in a[6:]], dt, sub_dt,
                                           ctypes.cast(None, _dp), ctypes.cast(None, _i64p), 1)
    else:
        st = ref_lib(lib).ref_diffusion_2d(_p(out), _p(a[0]), _p(a[1]), nx, ny, _p(a[2]), _p(a[3]), _p(a[4]),
                                        _p(a[5]), *types, *[_p(v) for v in a[6:]], dt, sub_dt)
    if st != 0:
        raise RuntimeError("diffusion_2d status %d" % st)
    return out


def assemble_2d(x, y, nx, ny, xl, xu, yl, yu, types, Dxx, Dyy, Dxy, Dyx, G, S, L, dt, lib="oracle"):
    n = nx * ny
    diag = np.zeros(9 * n)
    B0 = np.zeros(n)
    C0 = np.zeros(n)
    a = _local_args(x, y, nx, ny, xl, xu, yl, yu, Dxx, Dyy, Dxy, Dyx, G, S, L)
    f = oracle_lib().orc_assemble_2d if lib == "oracle" else ref_lib(lib).ref_assemble_2d
    st = f(nx, ny, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), _p(a[4]), _p(a[5]), *types, *[_p(v) for v in a[6:]], dt,
           _p(diag), _p(B0), _p(C0))
    if st != 0:
        raise RuntimeError("assemble status %d" % st)
    return diag.reshape(9, n), B0, C0


def pack_band(nx, ny, diag):
    n = nx * ny
    kl = nx + 1
    ldab = 3 * kl + 1
    ab = np.zeros(ldab * n)
    d = _c(diag)
    oracle_lib().orc_pack_band(nx, ny, _p(d), _p(ab))
    return ab.reshape(n, ldab).T.copy(order="F")  # (ldab, n) column-major like LAPACK


def dgbtf2(ab, kl, ku):
    """Returns (lu (ldab, n) Fortran-ordered, ipiv 1-based int64, info)."""
    lu = np.asfortranarray(ab, dtype=np.float64).copy(order="F")
    ldab, n = lu.shape
    ipiv = np.zeros(n, dtype=np.int64)
    info = oracle_lib().orc_dgbtf2(n, n, kl, ku, lu.ctypes.data_as(_dp), ldab, ipiv.ctypes.data_as(_i64p))
    return lu, ipiv, info


def dgbtrs(lu, kl, ku, b, ipiv):
    lu = np.asfortranarray(lu, dtype=np.float64)
    ldab, n = lu.shape
    x = _c(b).copy()
    piv = np.ascontiguousarray(ipiv, dtype=np.int64)
    oracle_lib().orc_dgbtrs(n, kl, ku, lu.ctypes.data_as(_dp), ldab, piv.ctypes.data_as(_i64p), _p(x))
    return x


def tridag(a, b, c, r, lib="oracle"):
    a, b, c, r = (_c(v) for v in (a, b, c, r))
    u = np.zeros_like(r)
    f = oracle_lib().orc_tridag if lib == "oracle" else ref_lib(lib).ref_tridag
    f(_p(a), _p(b), _p(c), _p(r), _p(u), len(r))
    return u


def diffusion_1d(psd, x, n, lbc, ubc, lt, ut, D, G, S, L, dt, lib="oracle"):
    out = _c(psd).copy()
    a = [_c(v) for v in (x, D, G, S, L)]
    if lib == "oracle":
        st = oracle_lib().orc_diffusion_1d(_p(out), _p(a[0]), n, lbc, ubc, lt, ut, _p(a[1]), _p(a[2]), _p(a[3]),
                                           _p(a[4]), dt)
    else:
        st = ref_lib(lib).ref_diffusion_1d(_p(out), _p(a[0]), n, lbc, ubc, lt, ut, _p(a[1]), _p(a[2]), _p(a[3]),
                                        _p(a[4]), dt)
    if st != 0:
        raise RuntimeError("diffusion_1d status %d" % st)
    return out


def quickest6(psd, x, n, lbc, ubc, lt, ut, ux, dt, lib="oracle"):
    out = _c(psd).copy()
    xx, uu = _c(x), _c(ux)
    if lib == "oracle":
        st = oracle_lib().orc_quickest6(_p(out), _p(xx), n, lbc, ubc, lt, ut, _p(uu), dt)
    else:
        st = ref_lib(lib).ref_quickest6(_p(out), _p(xx), n, lbc, ubc, lt, ut, _p(uu), dt)
    if st != 0:
        raise RuntimeError("quickest6 status %d" % st)
    return out


def convection_2d(psd, P, R, nP, nR, Pl, Pu, Rl, Ru, types, VP, VR, S, L, G, dt, lib="oracle"):
    out = _c(psd).copy()
    a = [_c(v) for v in (P, R, Pl, Pu, Rl, Ru, VP, VR, S, L, G)]
    f = oracle_lib().orc_convection_2d if lib == "oracle" else ref_lib(lib).ref_convection_2d
    st = f(_p(out), _p(a[0]), _p(a[1]), nP, nR, _p(a[2]), _p(a[3]), _p(a[4]), _p(a[5]), *types,
           *[_p(v) for v in a[6:]], dt)
    if st != 0:
        raise RuntimeError("convection_2d status %d" % st)
    return out


def convection_2d_ex(psd, lost, P, R, nP, nR, Pl, Pu, Rl, Ru, types, VP, VR, S, L, G, dt, fast=False,
                     lib="oracle"):
    """Convection_2D with the reference's build options: lost (None: off)
    is the SAVE_PSD_LOST_CONV accumulator, fast the FAST_CONVECTION rule.
    lib "oracle" is the restatement; "lost" / "fast" are the reference built
    with -DSAVE_PSD_LOST_CONV / -DFAST_CONVECTION. Returns (psd, lost)."""
    out = _c(psd).copy()
    lo = None if lost is None else _c(lost).copy()
    a = [_c(v) for v in (P, R, Pl, Pu, Rl, Ru, VP, VR, S, L, G)]
    if lib == "oracle":
        st = oracle_lib().orc_convection_2d_ex(_p(out), _p(lo) if lo is not None else ctypes.cast(None, _dp),
                                               _p(a[0]), _p(a[1]), nP, nR, _p(a[2]), _p(a[3]), _p(a[4]), _p(a[5]),
                                               *types, *[_p(v) for v in a[6:]], dt, 1 if fast else 0)
    elif lib == "lost":
        assert lo is not None and not fast
        st = ref_lib("lost").ref_convection_2d_lost(_p(out), _p(lo), _p(a[0]), _p(a[1]), nP, nR, _p(a[2]),
                                                    _p(a[3]), _p(a[4]), _p(a[5]), *types, *[_p(v) for v in a[6:]],
                                                    dt)
    else:
        assert lib == "fast" and lo is None
        st = ref_lib("fast").ref_convection_2d(_p(out), _p(a[0]), _p(a[1]), nP, nR, _p(a[2]), _p(a[3]), _p(a[4]),
                                               _p(a[5]), *types, *[_p(v) for v in a[6:]], dt)
    if st != 0:
        raise RuntimeError("convection_2d status %d" % st)
    return out, lo


def grid_struct(w):
    """Grid structure pointing at the arrays of a workload object (kept alive
    by the workload)."""
    g = Grid()
    w.fill_struct(g, lambda name, a: a.ctypes.data)
    return g


def step(w, psd, nthreads=0, run_convection=1, run_radial=1, run_local=1, lib="oracle"):
    """One full operator-split step of workload w on a copy of psd."""
    out = _c(psd).copy()
    g = grid_struct(w)
    if lib == "oracle":
        st = oracle_lib().orc_step(ctypes.byref(g), _p(out), nthreads, run_convection, run_radial, run_local)
    else:
        st = ref_lib(lib).ref_step_sample(ctypes.byref(g), _p(out), nthreads, run_convection, run_radial, run_local,
                                       0, -1, 0, -1, 0, -1)
    if st != 0:
        raise RuntimeError("step status %d" % st)
    return out


def ref_step_sample(w, psd, nthreads, q, l, nn):
    """Reference step restricted to (V,K) planes [0,q), lines [0,l), nodes
    [0,nn) (timing sample); psd is updated in place."""
    g = grid_struct(w)
    return ref_lib().ref_step_sample(ctypes.byref(g), _p(psd), nthreads, 1, 1, 1, 0, q, 0, l, 0, nn)


def cubic1d(x_old, values, x_new, lib="oracle"):
    """Cubic1D (src/Interpolation.cpp:25-51): values on x_old interpolated
    to x_new (same length); returns (status, result)."""
    xo, xn = _c(x_old), _c(x_new)
    out = _c(values).copy()
    assert xo.size == xn.size == out.size
    f = oracle_lib().orc_cubic1d if lib == "oracle" else ref_lib(lib).ref_cubic1d
    st = f(out.size, _p(xo), _p(out), _p(xn))
    return st, out


def remap(psd, l_old, l_new, lib="oracle"):
    """The remapping loop (src/VERB4D_Solver.cpp:574-606) on a copy of psd
    [P][R][V][K]; returns (status, result)."""
    out = _c(psd).copy()
    lo, ln = _c(l_old), _c(l_new)
    nP, nR, nV, nK = out.shape
    assert lo.shape == ln.shape == out.shape
    f = oracle_lib().orc_remap if lib == "oracle" else ref_lib(lib).ref_remap
    st = f(_p(out), _p(lo), _p(ln), nP, nR, nV, nK)
    return st, out
