"""CPU restatement of the reference time loop's pieces around the operators
and of UpdatableMatrix::update. TEST INFRASTRUCTURE ONLY: the checker of
verb4d.timeloop (the product drives the same sequence on the device).

  UpdatableArray.update   src/UpdatableMatrix.cpp:601-827 (file queue
                          :636-663, scale queue :674-713, limit queues
                          :715-800, MatrixLimit :269-300)
  step                    src/VERB4D_Solver.cpp:557-619 (remap), 621 (clamp), 633-740
                          (updates), 747-753 (stability check), 851-978
                          (convection, Vl/Vu from convection :965-975),
                          980 / 1072 (clamps), 990-1216 (radial, local),
                          1218-1238 (clamp or negative count)
"""
from collections import deque

import numpy as np

import oracle as O


class UpdatableArray:
    """files: [(t, array | path | "-")] (paths read by `loader` when due);
    scales: [(t, s)], `scale` the coefficient before them; limits: {ax: [(t,
    from, to)]} or limit_from / limit_to {ax: [(t, v)]}, limit_init {ax:
    (from, to)} constants; grids: the coordinate arrays of the limited axes."""

    def __init__(self, arr, files=(), scales=(), limits=None, grids=(), scale=1.0, limit_from=None, limit_to=None,
                 limit_init=None, loader=None):
        self.original = np.array(arr, dtype=np.float64, copy=True)
        self.current = self.original.copy()
        self.files = deque((float(t), a if isinstance(a, str) else np.asarray(a, dtype=np.float64)) for t, a in files)
        self.loader = loader
        self.scales = deque((float(t), float(s)) for t, s in scales)
        self.from_q = {ax: deque((float(t), float(v)) for t, v in q) for ax, q in (limit_from or {}).items()}
        self.to_q = {ax: deque((float(t), float(v)) for t, v in q) for ax, q in (limit_to or {}).items()}
        for ax, q in (limits or {}).items():
            self.from_q.setdefault(ax, deque()).extend((float(t), float(f)) for t, f, _ in q)
            self.to_q.setdefault(ax, deque()).extend((float(t), float(to)) for t, _, to in q)
        self.grids = grids
        self.last_update_time = -1.0
        self.latest_scale = float(scale)
        self.rng = {ax: [-1e99, 1e99] for ax in range(4)}
        for ax, (f, to) in (limit_init or {}).items():
            self.rng[ax] = [float(f), float(to)]
        self.limited = set(self.from_q) | set(self.to_q) | set(limit_init or {})

    def update(self, t):
        updated = False
        data, ut = None, None
        while self.files and t >= self.files[0][0]:
            ut, data = self.files.popleft()
        if data is not None and not (isinstance(data, str) and data in ("", "-")) and t >= ut \
                and self.last_update_time < ut:
            if isinstance(data, str):
                data = self.loader(data)
            self.original = np.asarray(data, dtype=np.float64).reshape(self.original.shape).copy()
            self.last_update_time = t
            updated = True
        cur = self.original.copy()
        while self.scales and t >= self.scales[0][0]:
            self.latest_scale = self.scales.popleft()[1]
            updated = True
        if self.latest_scale != 1:
            cur = cur * self.latest_scale
        for qs, k in ((self.to_q, 1), (self.from_q, 0)):
            for ax, q in qs.items():
                while q and t >= q[0][0]:
                    self.rng[ax][k] = q.popleft()[1]
                    updated = True
        # MatrixLimit: zero where any limited coordinate is out of range
        if self.limited:
            out = np.zeros(cur.shape, dtype=bool)
            for ax in self.limited:
                q = self.grids[ax]
                out |= (q < self.rng[ax][0]) | (q > self.rng[ax][1])
            cur = np.where(out, 0.0, cur)
        self.current = cur
        return updated


def stability_min(w):
    """(DVV.times(DKK) - DVK.times(DVK)).min(), NaN entries passed over."""
    v = w.f["DVV"] * w.f["DKK"] - w.f["DVK"] * w.f["DVK"]
    v = v[~np.isnan(v)]
    return float(v.min()) if v.size else float("inf")


def step(w, psd, fields, t, positive_psd=False, vl_from_convection=False, vu_from_convection=False, threads=8,
         remap=None):
    """One step of the loop on the workload w (its arrays are updated in
    place from `fields` {name: UpdatableArray}); returns (psd, info).
    remap: {"lstar": UpdatableArray, "copy": previous L*, "jacobians":
    {name: UpdatableArray}} (updated in place), src/VERB4D_Solver.cpp:557-619."""
    floor = 1e-21
    x = np.array(psd, copy=True)
    remapped = False
    if remap is not None and remap["lstar"].update(t):
        for n, f in remap["jacobians"].items():
            f.update(t)
            w.f[n][...] = f.current
        st, x = O.remap(x, remap["copy"], remap["lstar"].current)
        assert st == 0, st
        remap["copy"] = remap["lstar"].current.copy()
        remapped = True
    if positive_psd:
        x = np.where(x < floor, floor, x)
    updated = {}
    for n, f in fields.items():
        updated[n] = f.update(t)
        tgt = w.f if n in w.f else w.bc
        tgt[n][...] = f.current
    m = stability_min(w)
    if m < 0:
        return None, {"unstable": True, "stability_min": m}
    x = O.step(w, x, threads, 1, 0, 0)
    if vl_from_convection and w.types["Vl"] == 1:
        w.bc["Vl"] = np.ascontiguousarray(x[:, :, 0, :])
    if vu_from_convection and w.types["Vu"] == 1:
        w.bc["Vu"] = np.ascontiguousarray(x[:, :, w.nV - 1, :])
    if positive_psd:
        x = np.where(x < floor, floor, x)
    x = O.step(w, x, threads, 0, 1, 0)
    if positive_psd:
        x = np.where(x < floor, floor, x)
    x = O.step(w, x, threads, 0, 0, 1)
    neg = 0
    if positive_psd:
        x = np.where(x < floor, floor, x)
    else:
        neg = int((x < 0).sum())
    return x, {"updated": updated, "negative_points": neg, "stability_min": m, "remapped": remapped}
