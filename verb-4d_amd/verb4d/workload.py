"""Synthetic, seeded inputs for the operator-split step.

Grids and coefficient laws of the benchmark configurations:
  * P (MLT): closed grid linspace(0, 2*pi, nP); point nP-1 duplicates point 0,
    the periodic convention QUICKEST6 assumes
    (src/Convection_1D_ULTIMATE_QUICKEST6.cpp:97-100).
  * R = L: linspace(1.5, 6.6, nR) (dipole field).
  * E: logspace(-2, 1, nV) MeV; alpha: linspace(alpha_lc(L), pi/2, nK) per L,
    with the dipole bounce loss cone sin^2(alpha_lc) = 1 / (L^3 sqrt(4 - 3/L)).
  * PSD0 = exp(-(L-4.5)^2/0.5) * exp(-E/0.3) * sin(alpha).
  * Daa, DEE log-uniform in [1e-7, 1e-3] /s, DaE = 0.5 sqrt(Daa DEE) U(-1,1),
    scaled by 10^((Kp-Kp0)/4) (x10 for the Kp 2 -> 6 step); mapped to the
    reference arrays DKK = Daa, DVV = DEE E^2, DVK = DaE E, so that
    DVV*DKK - DVK^2 >= 0.75 E^2 Daa DEE > 0 (src/VERB4D_Solver.cpp:747).
  * D_LL = 10^(0.506 Kp - 9.325) L^10 /day, in /s.
  * VP, VR: Volland-Stern E x B drift (gamma = 2, Maynard-Chen A0(Kp)) plus
    corotation in a dipole, in rad/s and L/s.
  * G_local = p^2 sin(a) cos(a) (1.30 - 0.56 sin(a)), G_radial = 1/L^2,
    G_conv = 1; Losses_local = -1e-6 /s (the reference stores losses as
    negative rates: A0 -= L, exp(L dt)); other sources and losses zero.
  * BCs: P periodic; R (= L) Dirichlet from PSD0; alpha-lower Dirichlet 0
    (loss cone), alpha-upper zero derivative; E Dirichlet from PSD0.

Random numbers come from a counter-based SplitMix64 of
(seed, field id, flat index) -> U[0,1) = (z >> 11) * 2^-53, so any subset of
the grid can be regenerated independently.
"""
import math

import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15

FIELD_DAA = 1
FIELD_DEE = 2
FIELD_DAE = 3
FIELD_SRC = 4
FIELD_LCONV = 5
FIELD_LRAD = 6
FIELD_LLOC = 7
FIELD_PSD = 8


def splitmix_u01(seed, field_id, n, offset=0):
    """U[0,1) doubles for flat indices offset .. offset+n-1."""
    base = ((int(seed) * GOLDEN) & M64) ^ ((int(field_id) & 0xFFFFFFFF) << 32)
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(base) + (idx + np.uint64(1)) * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def loss_cone(L):
    s2 = 1.0 / (L ** 3 * np.sqrt(4.0 - 3.0 / L))
    return np.arcsin(np.sqrt(s2))


def volland_stern(P, L, kp):
    """(VP [rad/s], VR [L/s]) of the E x B drift plus corotation in a dipole."""
    RE = 6.371e6
    B0 = 3.11e-5
    omega = 2.0 * math.pi / 86400.0
    a0 = 0.045 / (1.0 - 0.159 * kp + 0.0093 * kp * kp) ** 3  # kV / RE^2
    a0_si = a0 * 1e3 / (RE * RE)
    VP = omega + 2.0 * a0_si * L ** 3 * np.sin(P) / B0
    VR = -a0_si * L ** 4 * np.cos(P) / B0
    return VP, VR


def dll(L, kp):
    return 10.0 ** (0.506 * kp - 9.325) * L ** 10 / 86400.0


class Workload:
    """All arrays of one configuration ([P][R][V][K] float64 unless noted)."""

    def __init__(self, nP, nR, nV, nK, seed, kp=3.0, kp0=None, dt=300.0, sub_dt=None, rich=False, stiff=False,
                 kp_by_mlt=None):
        self.nP, self.nR, self.nV, self.nK = nP, nR, nV, nK
        self.seed = seed
        self.kp0 = kp if kp0 is None else kp0
        self.dt = float(dt)
        self.sub_dt = float(dt if sub_dt is None else sub_dt)
        self.rich = rich
        shape = (nP, nR, nV, nK)
        self.shape = shape
        P = np.linspace(0.0, 2.0 * math.pi, nP)
        L = np.linspace(1.5, 6.6, nR)
        E = np.logspace(-2.0, 1.0, nV)
        alpha = np.empty((nR, nK))
        for r in range(nR):
            alpha[r] = np.linspace(loss_cone(L[r]), math.pi / 2, nK)
        self.P1, self.L1, self.E1, self.alpha = P, L, E, alpha
        Pg = np.broadcast_to(P[:, None, None, None], shape)
        Lg = np.broadcast_to(L[None, :, None, None], shape)
        Eg = np.broadcast_to(E[None, None, :, None], shape)
        Ag = np.broadcast_to(alpha[None, :, None, :], shape)
        self.f = {}
        self.f["Pg"] = np.ascontiguousarray(Pg)
        self.f["Rg"] = np.ascontiguousarray(Lg)
        self.f["Lg"] = self.f["Rg"]
        self.f["Vg"] = np.ascontiguousarray(Eg)
        self.f["Kg"] = np.ascontiguousarray(Ag)
        self.psd0 = np.ascontiguousarray(np.exp(-((Lg - 4.5) ** 2) / 0.5) * np.exp(-Eg / 0.3) * np.sin(Ag))
        n = nP * nR * nV * nK
        if stiff:
            # Daa log-uniform in [1e-9, 1] /s, largest within ~5 degrees of
            # the loss cone (two decades lower far from it); DEE log-uniform
            # in [1e-8, 1e-2] /s
            alc = loss_cone(Lg)
            near = np.exp(-(((Ag - alc) / math.radians(5.0)) ** 2))
            self.daa = (10.0 ** (-9.0 + 9.0 * splitmix_u01(seed, FIELD_DAA, n))).reshape(shape) * (
                0.01 + 0.99 * near)
            self.dee = (10.0 ** (-8.0 + 6.0 * splitmix_u01(seed, FIELD_DEE, n))).reshape(shape)
        else:
            self.daa = (10.0 ** (-7.0 + 4.0 * splitmix_u01(seed, FIELD_DAA, n))).reshape(shape)
            self.dee = (10.0 ** (-7.0 + 4.0 * splitmix_u01(seed, FIELD_DEE, n))).reshape(shape)
        self.dae_u = (2.0 * splitmix_u01(seed, FIELD_DAE, n) - 1.0).reshape(shape)
        p2 = Eg * (Eg + 2.0 * 0.511)
        sa, ca = np.sin(Ag), np.cos(Ag)
        self.f["G_local"] = np.ascontiguousarray(p2 * sa * ca * (1.30 - 0.56 * sa))
        self.f["G_radial"] = np.ascontiguousarray(1.0 / Lg ** 2)
        self.f["G_conv"] = np.ones(shape)
        if rich:
            # exercise every term: small positive sources, small loss rates
            self.f["sources"] = (1e-9 * splitmix_u01(seed, FIELD_SRC, n)).reshape(shape) * self.psd0.max()
            self.f["losses_conv"] = -(1e-5 * splitmix_u01(seed, FIELD_LCONV, n)).reshape(shape)
            self.f["losses_radial"] = -(1e-5 * splitmix_u01(seed, FIELD_LRAD, n)).reshape(shape)
            self.f["losses_local"] = -(1e-5 * splitmix_u01(seed, FIELD_LLOC, n)).reshape(shape)
        else:
            self.f["sources"] = np.zeros(shape)
            self.f["losses_conv"] = np.zeros(shape)
            self.f["losses_radial"] = np.zeros(shape)
            self.f["losses_local"] = np.full(shape, -1e-6)
        # boundaries
        self.bc = {}
        self.bc["Pl"] = np.full((nR, nV, nK), -1e99)
        self.bc["Pu"] = np.full((nR, nV, nK), -1e99)
        self.bc["Rl"] = np.ascontiguousarray(self.psd0[:, 0, :, :])
        self.bc["Ru"] = np.ascontiguousarray(self.psd0[:, nR - 1, :, :])
        self.bc["Ll"] = self.bc["Rl"]
        self.bc["Lu"] = self.bc["Ru"]
        self.bc["Vl"] = np.ascontiguousarray(self.psd0[:, :, 0, :])
        self.bc["Vu"] = np.ascontiguousarray(self.psd0[:, :, nV - 1, :])
        self.bc["Kl"] = np.zeros((nP, nR, nV))
        self.bc["Ku"] = np.zeros((nP, nR, nV))
        self.types = dict(Pl=0, Pu=0, Rl=1, Ru=1, Ll=1, Lu=1, Vl=1, Vu=1, Kl=1, Ku=2)
        self.radial_sources = 0.0
        self.local_sources = 1.0
        self.include_boundary = 1
        # optional Kp per MLT sector (array of nP): mixed regimes across MLT
        self.kp_by_mlt = None if kp_by_mlt is None else np.asarray(kp_by_mlt, dtype=np.float64)
        self.set_kp(kp)

    def set_kp(self, kp):
        """Coefficients that follow Kp: local D's, D_LL and the drift field."""
        self.kp = float(kp)
        if self.kp_by_mlt is not None:
            kpp = self.kp_by_mlt[:, None, None, None]
            s = 10.0 ** ((kpp - self.kp0) / 4.0)
        else:
            kpp = self.kp
            s = 10.0 ** ((self.kp - self.kp0) / 4.0)
        Eg = self.f["Vg"]
        self.f["DKK"] = np.ascontiguousarray(self.daa * s)
        self.f["DVV"] = np.ascontiguousarray(self.dee * s * Eg ** 2)
        self.f["DVK"] = np.ascontiguousarray(0.5 * np.sqrt(self.daa * s * self.dee * s) * self.dae_u * Eg)
        self.f["DLL"] = np.ascontiguousarray(dll(self.f["Lg"], kpp))
        VP, VR = volland_stern(self.f["Pg"], self.f["Rg"], kpp)
        self.f["VP"] = np.ascontiguousarray(VP)
        self.f["VR"] = np.ascontiguousarray(VR)

    @property
    def points(self):
        return self.nP * self.nR * self.nV * self.nK

    def fill_struct(self, g, ptr_of):
        """Fill a verb_grid-layout ctypes Structure; ptr_of(array) -> address."""
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


def lstar(w, compression):
    """Synthetic L* on the grid, [P][R][V][K]: the dipole L lowered by
    `compression` x (L / L_max)^2 at noon (MLT angle 0), not at all at
    midnight, more for equatorial than for field-aligned pitch angles (drift
    shell splitting). Increasing along R while compression < 1/3;
    compression 0 gives L* = L."""
    P, L, A = w.f["Pg"], w.f["Rg"], w.f["Kg"]
    Lmax = float(L.max())
    day = 0.5 * (1.0 + np.cos(P))
    return np.ascontiguousarray(L * (1.0 - compression * (L / Lmax) ** 2 * day * (0.6 + 0.4 * np.sin(A))))


# Benchmark configurations (the bench's timed workload is PRODUCTION).
def config0(seed=1, rich=False):
    """10x24x31x31 (R x MLT x E x alpha), Kp = 3, dt = 300 s, 12 steps."""
    return Workload(nP=24, nR=10, nV=31, nK=31, seed=seed, kp=3.0, dt=300.0, rich=rich)


def config1(seed=1001, rich=False):
    """29x24x101x91, Kp 2 -> 6 at step 50, dt = 300 s, 100 steps."""
    return Workload(nP=24, nR=29, nV=101, nK=91, seed=seed, kp=2.0, kp0=2.0, dt=300.0, rich=rich)


FIELD_BAND = 9


def config2(seed=2001):
    """Uneven systems on the production grid: returns (workload, k0) with k0
    [P][R] the first active pitch-angle column of each node. The active
    count follows the bounce loss cone, 40 cells at R = 1.5 rising linearly
    to 91 at R = 6.6, plus an MLT-dependent drift-loss band of -10..+10
    cells drawn from the seed (clipped to 40..91), so systems have
    101*40 = 4040 to 101*91 = 9191 unknowns. Kp is 2 on the night side
    (MLT sectors 0..11) and 6 on the day side (12..23)."""
    nP, nR, nV, nK = 24, 29, 101, 91
    kp = np.where(np.arange(nP) < nP // 2, 2.0, 6.0)
    w = Workload(nP=nP, nR=nR, nV=nV, nK=nK, seed=seed, kp=2.0, kp0=2.0, dt=300.0, kp_by_mlt=kp)
    base = np.round(40.0 + (nK - 40.0) * (w.L1 - 1.5) / (6.6 - 1.5))
    band = np.floor(21.0 * splitmix_u01(seed, FIELD_BAND, nP * nR)).reshape(nP, nR) - 10.0
    active = np.clip(base[None, :] + band, 40, nK).astype(np.int32)
    return w, (nK - active).astype(np.int32)


def config3(seed=3001, nP=24, nR=29):
    """Stiff mixed-precision case: 29x24x151x121, Daa log-uniform in
    1e-9..1 /s peaking near the loss cone, DEE log-uniform in 1e-8..1e-2 /s,
    dt = 600 s (kl = ku = 152 in the reference's energy-fast ordering)."""
    return Workload(nP=nP, nR=nR, nV=151, nK=121, seed=seed, kp=2.0, kp0=2.0, dt=600.0, stiff=True)


def kp_schedule_config1(step):
    return 2.0 if step < 50 else 6.0
