"""The reference's time loop around the three operators, driven on the
device-resident grid (src/VERB4D_Solver.cpp:539-1282):

  remap (when L* changes) -> clamp -> coefficient updates -> boundary
  updates -> stability check ->
  convection (-> V boundaries from the PSD) -> clamp -> radial -> clamp ->
  local (LU factors recomputed when a local coefficient changed) ->
  clamp or negative-point count

UpdatableField mirrors UpdatableMatrix::update (src/UpdatableMatrix.cpp:
601-827): the host keeps the time queues (replacement arrays as a .lst list
would load them, scaling coefficients, per-axis limits) and the device
array is recomputed by verb_field_update (scaling, then limiting) whenever
an update is due. The loop's arithmetic all runs in the library.
"""
import ctypes
from collections import deque

import numpy as np

from . import FieldUpdateParams, StepOptions, VerbError, lib


class UpdatableField:
    """One device field of a DeviceGrid with UpdatableMatrix semantics.

    files:  [(time, array | path | "-")] replacement originals (the .lst
            queue: the last entry due at the current time is loaded, once
            per update time, src/UpdatableMatrix.cpp:636-663; a path is read
            with `loader` when it comes due, "-" loads nothing)
    scales: [(time, coefficient)] (the scale queue, :674-713); `scale` is the
            coefficient before the first entry (a constant scale field)
    limits: {axis: [(time, from, to)]}, or separate `limit_from` /
            `limit_to` {axis: [(time, value)]} queues, with axis in 0..3 of
            `axes` (the Q*_from / Q*_to queues, :715-800); `limit_init`
            {axis: (from, to)} are constant limits
    axes:   names of the DeviceGrid fields holding each point's coordinate
            along each axis (4-D fields: Pg, Rg, Vg, Kg)
    tensor: the device array to update when it is not one of the grid's
            fields (L* for the remapping)."""

    def __init__(self, dg, name, files=(), scales=(), limits=None, axes=("Pg", "Rg", "Vg", "Kg"), tensor=None,
                 scale=1.0, limit_from=None, limit_to=None, limit_init=None, loader=None):
        import torch
        self.dg = dg
        self.name = name
        self.cur = dg.t[name][1] if tensor is None else tensor
        self.original = self.cur.clone()
        torch.cuda.synchronize()
        self.files = deque((float(t), a if isinstance(a, str) else np.ascontiguousarray(a, dtype=np.float64))
                           for t, a in files)
        self.loader = loader
        self.scales = deque((float(t), float(v)) for t, v in scales)
        self.from_q = {ax: deque((float(t), float(v)) for t, v in q) for ax, q in (limit_from or {}).items()}
        self.to_q = {ax: deque((float(t), float(v)) for t, v in q) for ax, q in (limit_to or {}).items()}
        for ax, q in (limits or {}).items():
            self.from_q.setdefault(ax, deque()).extend((float(t), float(f)) for t, f, _ in q)
            self.to_q.setdefault(ax, deque()).extend((float(t), float(to)) for t, _, to in q)
        self.axes = axes
        self.last_update_time = -1.0
        self.scale = float(scale)
        self.range = {ax: [-1e99, 1e99] for ax in range(4)}
        for ax, (f, to) in (limit_init or {}).items():
            self.range[ax] = [float(f), float(to)]
        self.limited = set(self.from_q) | set(self.to_q) | set(limit_init or {})

    @classmethod
    def from_spec(cls, dg, name, spec, grids, axes=("Pg", "Rg", "Vg", "Kg"), tensor=None):
        """From verb4d.formats.field_spec (a .tab line): .lst entries are
        read, grid-checked against `grids` (host arrays), when due."""
        from . import formats
        f = cls(dg, name, files=spec["files"], scales=spec["scales"], axes=axes, tensor=tensor,
                scale=spec["scale"], limit_from=spec["limit_from"], limit_to=spec["limit_to"],
                limit_init=spec["limit_init"], loader=lambda p: formats.read_any(p, grids))
        return f

    def update(self, time):
        """Recomputes the device array for `time`; True when something was
        due (a new original, a scale or a limit entry)."""
        import torch
        updated = False
        data, utime = None, None
        while self.files and time >= self.files[0][0]:
            utime, data = self.files.popleft()
        if data is not None and not (isinstance(data, str) and data in ("", "-")) and time >= utime \
                and self.last_update_time < utime:
            if isinstance(data, str):
                data = self.loader(data)
            self.dg.ctx.synchronize()
            self.original.copy_(torch.from_numpy(np.ascontiguousarray(data, dtype=np.float64).reshape(
                self.original.shape)))
            torch.cuda.synchronize()
            self.last_update_time = time
            updated = True
        while self.scales and time >= self.scales[0][0]:
            self.scale = self.scales.popleft()[1]
            updated = True
        for qs, k in ((self.to_q, 1), (self.from_q, 0)):
            for ax, q in qs.items():
                while q and time >= q[0][0]:
                    self.range[ax][k] = q.popleft()[1]
                    updated = True
        p = FieldUpdateParams()
        p.scale = self.scale
        for ax in range(4):
            f, to = self.range[ax]
            p.from_[ax], p.to[ax] = f, to
            p.grid[ax] = self.dg.t[self.axes[ax]][1].data_ptr() if (ax in self.limited and ax < len(self.axes)) \
                else None
        self.dg.ctx.check(lib().verb_field_update(self.dg.ctx.handle, self.cur.data_ptr(), self.original.data_ptr(),
                                                  self.cur.numel(), ctypes.byref(p)))
        return updated


class TimeLoop:
    """The per-step sequence of the reference's main loop on a DeviceGrid.

    fields: {name: UpdatableField} of coefficients (DKK, DVV, DVK,
            losses_local, DLL, VP, VR, sources, ...) and boundaries (Vu, ...)
    positive_psd, vl_from_convection, vu_from_convection: the reference's
            switches (src/VERB4D_Solver.cpp:621, 965-975, 1218-1221)
    lu_caching: keep the band factors between steps and recompute them only
            when losses_local, DVV, DVK or DKK changed (:1091-1096); the
            first step always factors (the cache starts empty).
    lstar:  an UpdatableField over the L* array ([P][R][V][K], `tensor=`):
            when it updates, the `jacobians` fields (G_local, G_radial,
            G_conv) update too and the PSD is remapped from the previous L*
            to the new one, before anything else of the step (:557-619)."""

    LOCAL = ("losses_local", "DVV", "DVK", "DKK")

    def __init__(self, dg, fields=None, positive_psd=False, vl_from_convection=False, vu_from_convection=False,
                 lu_caching=False, time_first=0.0, dt_days=None, lstar=None, jacobians=None):
        import torch
        self.dg = dg
        self.fields = dict(fields or {})
        self.lstar = lstar
        self.jacobians = dict(jacobians or {})
        self.lstar_copy = None
        if lstar is not None:
            dg.ctx.synchronize()
            self.lstar_copy = lstar.cur.clone()
            torch.cuda.synchronize()
        self.positive_psd = positive_psd
        w = dg.w
        self.vl = torch.empty((w.nP, w.nR, w.nK), dtype=torch.float64, device="cuda") if vl_from_convection else None
        self.vu = torch.empty((w.nP, w.nR, w.nK), dtype=torch.float64, device="cuda") if vu_from_convection else None
        self.lu_caching = lu_caching
        if lu_caching and dg.lu_cache is None:
            dg.alloc_lu_cache()
        self.factored = False
        self.time_first = time_first
        self.dt_days = w.dt / 86400.0 if dt_days is None else dt_days
        self.last = {}

    def step(self, it):
        """Step `it`: returns {'updated': {name: bool}, 'negative_points': n,
        'recompute_lu': bool, 'stability_min': value, 'time': days,
        'remapped': bool}."""
        import torch
        dg = self.dg
        time = self.time_first + it * self.dt_days
        remapped = False
        if self.lstar is not None and self.lstar.update(time):
            for f in self.jacobians.values():
                f.update(time)
            dg.remap(self.lstar_copy, self.lstar.cur)
            dg.ctx.synchronize()
            self.lstar_copy.copy_(self.lstar.cur)
            torch.cuda.synchronize()
            remapped = True
        updated = {n: f.update(time) for n, f in self.fields.items()}
        m = ctypes.c_double()
        rc = lib().verb_check_stability(dg.ctx.handle, ctypes.byref(dg.grid), ctypes.byref(m))
        if rc != 0:
            raise VerbError(rc, "Fatal error: (DVV*DKK - DVK*DVK) = %g < 0. The computation would be unstable."
                            % m.value)
        recompute = (not self.factored) or any(updated.get(n, False) for n in self.LOCAL)
        cnt = ctypes.c_longlong(-1)
        opt = StepOptions(1 if self.positive_psd else 0, self.vl.data_ptr() if self.vl is not None else None,
                          ctypes.pointer(cnt), None, 0, self.vu.data_ptr() if self.vu is not None else None)
        lu = dg.lu_cache.data_ptr() if self.lu_caching else None
        piv = dg.ipiv_cache.data_ptr() if self.lu_caching else None
        dg.ctx.check(lib().verb_step_ex(dg.ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(), 1, 1, 1, lu, piv,
                                        1 if recompute else 0, ctypes.byref(opt)))
        self.factored = self.factored or self.lu_caching
        # the boundaries taken from convection stay in force until the next
        # update replaces them (the reference keeps Vl_BC / Vu_BC)
        if self.vl is not None and dg.w.types["Vl"] == 1:
            dg.grid.Vl = self.vl.data_ptr()
        if self.vu is not None and dg.w.types["Vu"] == 1:
            dg.grid.Vu = self.vu.data_ptr()
        self.last = {"updated": updated, "negative_points": int(cnt.value), "recompute_lu": recompute,
                     "stability_min": float(m.value), "time": time, "remapped": remapped}
        return self.last


__all__ = ["UpdatableField", "TimeLoop"]
