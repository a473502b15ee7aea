"""Device-resident full-grid state for the batched entries.

Torch supplies HBM allocations only; every operator runs through the C-ABI
(``verb_step`` and the three batched entries of include/verb4d_hip.h) on the
context's own HIP stream.
"""
import ctypes

import numpy as np

from . import Context, GmresParams, GmresStats, MixedParams, MixedStats, StepOptions, Tuning, VerbGrid, lib


class DeviceGrid:
    """All fields of a workload in HBM plus the ``verb_grid`` that points at
    them. ``psd`` is a float64 device tensor [P][R][V][K]."""

    def __init__(self, workload, ctx=None, device="cuda"):
        import torch

        self.torch = torch
        self.w = workload
        self.ctx = ctx or Context()
        self.device = device
        self.t = {}
        self.grid = VerbGrid()
        workload.fill_struct(self.grid, self._upload)
        self.psd = torch.from_numpy(np.ascontiguousarray(workload.psd0)).to(device)
        self.lu_cache = None
        self.ipiv_cache = None
        torch.cuda.synchronize()

    def _upload(self, name, arr):
        torch = self.torch
        key = id(arr)
        for k, (aid, tens) in self.t.items():
            if aid == key:
                return tens.data_ptr()
        tens = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(self.device)
        self.t[name] = (key, tens)
        return tens.data_ptr()

    def refresh(self, names):
        """Re-upload fields of the workload that changed on the host (e.g.
        after ``Workload.set_kp``)."""
        torch = self.torch
        # the library's stream may still be reading these fields
        self.ctx.synchronize()
        for n in names:
            src = self.w.f[n] if n in self.w.f else self.w.bc[n]
            _, tens = self.t[n]
            tens.copy_(torch.from_numpy(np.ascontiguousarray(src)))
        torch.cuda.synchronize()

    def set_psd(self, arr):
        # the copy runs on torch's stream: wait for the library's stream first,
        # or a step still in flight there could overwrite the new contents
        self.ctx.synchronize()
        self.psd.copy_(self.torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)))
        self.torch.cuda.synchronize()

    def psd_host(self):
        self.ctx.synchronize()
        return self.psd.cpu().numpy()

    def alloc_lu_cache(self):
        torch = self.torch
        nd = lib().verb_lu_cache_doubles(ctypes.byref(self.grid))
        ni = lib().verb_lu_cache_ints(ctypes.byref(self.grid))
        self.lu_cache = torch.empty(nd, dtype=torch.float64, device=self.device)
        self.ipiv_cache = torch.empty(ni, dtype=torch.int64, device=self.device)

    def step(self, run_convection=True, run_radial=True, run_local=True, use_cache=False, recompute_lu=True):
        lu = self.lu_cache.data_ptr() if use_cache else None
        piv = self.ipiv_cache.data_ptr() if use_cache else None
        rc = lib().verb_step(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(), int(run_convection),
                             int(run_radial), int(run_local), lu, piv, 1 if recompute_lu else 0)
        self.ctx.check(rc)

    def step_ex(self, positive_psd=False, vl_from_convection=None, count_negative=False, run_convection=True,
                run_radial=True, run_local=True, psd_lost=None, fast_convection=False):
        """verb_step_ex: the step with the time loop's positivity clamps,
        the V-lower boundary taken from the PSD after convection (into the
        given device tensor [P][R][K]), the count of negative points and the
        convection options (psd_lost: a device tensor like psd that
        accumulates the PSD the convection losses take). Returns the count
        (or None)."""
        cnt = ctypes.c_longlong(-1)
        opt = StepOptions(1 if positive_psd else 0,
                          vl_from_convection.data_ptr() if vl_from_convection is not None else None,
                          ctypes.pointer(cnt) if count_negative else None,
                          psd_lost.data_ptr() if psd_lost is not None else None, 1 if fast_convection else 0)
        rc = lib().verb_step_ex(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(), int(run_convection),
                                int(run_radial), int(run_local), None, None, 1, ctypes.byref(opt))
        self.ctx.check(rc)
        return cnt.value if count_negative else None

    def assembly(self):
        """The assembled local operator of every node (verb_local_assembly):
        (diag [nodes][9][n], B0 [nodes][n], C0 [nodes][n]) as host arrays,
        energy-fast unknowns."""
        torch = self.torch
        nodes = self.w.nP * self.w.nR
        n = self.w.nV * self.w.nK
        diag = torch.zeros((nodes, 9, n), dtype=torch.float64, device=self.device)
        b0 = torch.zeros((nodes, n), dtype=torch.float64, device=self.device)
        c0 = torch.zeros((nodes, n), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize()
        self.ctx.check(lib().verb_local_assembly(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                 diag.data_ptr(), b0.data_ptr(), c0.data_ptr()))
        self.ctx.synchronize()
        return diag.cpu().numpy(), b0.cpu().numpy(), c0.cpu().numpy()

    def convection(self):
        self.ctx.check(lib().verb_convection_batched(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr()))

    def radial(self):
        self.ctx.check(lib().verb_radial_diffusion_batched(self.ctx.handle, ctypes.byref(self.grid),
                                                           self.psd.data_ptr()))

    def da_radial(self, obs, covariance, dt, model_error, observation_error, p_index=0, model_matrix=None):
        """verb_da_radial: the Kalman update along L of every (V,K) line of P
        slice p_index; obs [P][R][V][K] and covariance [V][K][R][R] device
        tensors (covariance in/out, zero before the first call)."""
        import torch
        from . import DaParams
        p = DaParams(float(dt), float(model_error), float(observation_error), int(p_index))
        torch.cuda.current_stream().synchronize()
        self.ctx.check(lib().verb_da_radial(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                            obs.data_ptr(), covariance.data_ptr(),
                                            model_matrix.data_ptr() if model_matrix is not None else None,
                                            ctypes.byref(p)))

    def da_convection(self, obs, loss, covariance, model, dt, correlation_time, model_error, observation_error,
                      saturation=None, use_log=False, growing_Q_at_boundary=False, recompute_model=True,
                      plane_begin=0, plane_count=0, process_noise=None):
        """verb_da_convection: the Kalman filter of every (V,K) plane in
        plane_begin .. +plane_count (0: to the last); obs, loss, saturation
        [P][R][V][K] and covariance, model [V][K][N][N] device tensors
        (covariance zero before the first call; model written when
        recompute_model, read otherwise; process_noise, optional, the same
        for Q). The solver passes its Sources array as loss
        (src/VERB4D_Solver.cpp:1246)."""
        import torch
        from . import DaConvParams
        p = DaConvParams(float(dt), float(correlation_time), float(model_error), float(observation_error),
                         int(bool(use_log)), int(bool(growing_Q_at_boundary)), int(bool(recompute_model)),
                         int(plane_begin), int(plane_count),
                         process_noise.data_ptr() if process_noise is not None else None)
        torch.cuda.current_stream().synchronize()
        self.ctx.check(lib().verb_da_convection(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                obs.data_ptr(), loss.data_ptr(),
                                                saturation.data_ptr() if saturation is not None else None,
                                                covariance.data_ptr(), model.data_ptr(), ctypes.byref(p)))

    def remap(self, lstar_old, lstar_new):
        """verb_remap_batched: the PSD interpolated along R from lstar_old to
        lstar_new (device tensors [P][R][V][K], e.g. from torch's stream:
        finished before the library reads them)."""
        import torch
        w = self.w
        for t in (lstar_old, lstar_new):
            assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == self.psd.numel()
        torch.cuda.current_stream().synchronize()
        self.ctx.check(lib().verb_remap_batched(self.ctx.handle, self.psd.data_ptr(), lstar_old.data_ptr(),
                                                lstar_new.data_ptr(), w.nP, w.nR, w.nV, w.nK))

    def local(self, use_cache=False, recompute_lu=True):
        lu = self.lu_cache.data_ptr() if use_cache else None
        piv = self.ipiv_cache.data_ptr() if use_cache else None
        self.ctx.check(lib().verb_local_diffusion_batched(self.ctx.handle, ctypes.byref(self.grid),
                                                          self.psd.data_ptr(), lu, piv, 1 if recompute_lu else 0))

    def local_uneven(self, k0):
        """Local diffusion with per-node first active K column k0 [P][R]."""
        import numpy as np
        k = np.ascontiguousarray(k0, dtype=np.int32).ravel()
        self.ctx.check(lib().verb_local_diffusion_uneven(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                         k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None,
                                                         1))

    def step_uneven(self, k0, run_convection=True, run_radial=True, run_local=True):
        import numpy as np
        k = np.ascontiguousarray(k0, dtype=np.int32).ravel()
        self.ctx.check(lib().verb_step_uneven(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                              k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(run_convection),
                                              int(run_radial), int(run_local)))

    @staticmethod
    def gmres_params(tol=1e-11, restart=20, max_cycles=8, inner_tol=1e-5, max_iterations=0):
        return GmresParams(tol, restart, max_cycles, inner_tol, max_iterations)

    @staticmethod
    def gmres_params_capped(tol=1e-10, iterations=8, cycles=2, inner_tol=1e-6):
        """At most `iterations` inner iterations per node in all (split over
        `cycles` FP32 cycles, each closed by an FP64 residual), then the band
        LU for every node that has not met tol."""
        return GmresParams(tol, (iterations + cycles - 1) // cycles, cycles, inner_tol, iterations)

    def local_gmres(self, params=None, stats=True, x0=None):
        """Local diffusion by matrix-free GMRES with band-LU fallback, from
        the incoming PSD or from the device tensor x0 [P][R][V][K]; returns
        the stats dict (synchronises) or None."""
        p = params or self.gmres_params()
        st = GmresStats() if stats else None
        if x0 is None:
            rc = lib().verb_local_diffusion_gmres(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                  ctypes.byref(p), ctypes.byref(st) if st else None)
        else:
            rc = lib().verb_local_diffusion_gmres_x0(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                     x0.data_ptr(), ctypes.byref(p), ctypes.byref(st) if st else None)
        self.ctx.check(rc)
        return st.as_dict() if st else None

    def step_gmres(self, params=None, stats=False):
        p = params or self.gmres_params()
        st = GmresStats() if stats else None
        self.ctx.check(lib().verb_step_gmres(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                             ctypes.byref(p), ctypes.byref(st) if st else None))
        return st.as_dict() if st else None

    def local_mixed(self, tol=1e-12, max_sweeps=8, stats=True):
        """Local diffusion with FP32 band LU + FP64 iterative refinement;
        returns the stats dict (synchronises) or None."""
        p = MixedParams(tol, max_sweeps)
        st = MixedStats() if stats else None
        self.ctx.check(lib().verb_local_diffusion_mixed(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                                        ctypes.byref(p), ctypes.byref(st) if st else None))
        return st.as_dict() if st else None

    def step_mixed(self, tol=1e-12, max_sweeps=8, stats=False):
        p = MixedParams(tol, max_sweeps)
        st = MixedStats() if stats else None
        self.ctx.check(lib().verb_step_mixed(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                             ctypes.byref(p), ctypes.byref(st) if st else None))
        return st.as_dict() if st else None

    def autotune(self):
        """Times the launch shapes on this grid and keeps the fastest in the
        context; returns them as a dict (psd is left unchanged)."""
        t = Tuning()
        self.ctx.check(lib().verb_autotune(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(),
                                           ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, convection_tq=0, lu_rb=0, lu_cb=0, lu_persistent=0, convection_kernel=0):
        t = Tuning(convection_tq, lu_rb, lu_cb, lu_persistent, 0.0, 0.0, convection_kernel)
        self.ctx.check(lib().verb_set_tuning(self.ctx.handle, ctypes.byref(t)))

    def step_graph(self, k0=None):
        """The plain step (uneven pitch-angle extents k0 [P][R] when given)
        through a captured HIP graph: captured on the first call for this
        grid, replayed afterwards."""
        kp = None
        if k0 is not None:
            self._k0_keep = np.ascontiguousarray(k0, dtype=np.int32).ravel()
            kp = self._k0_keep.ctypes.data_as(ctypes.c_void_p)
        self.ctx.check(lib().verb_step_graph(self.ctx.handle, ctypes.byref(self.grid), self.psd.data_ptr(), kp))

    def check_status(self):
        return self.ctx.check_status()
