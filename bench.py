"""Full-grid operator-split step throughput on one MI355X.

A "step" is one pass of the hot path over the whole grid: convection
(Convection_2D over every (V,K) plane), radial diffusion (Diffusion_1D over
every (P,V,K) line) and local diffusion (Diffusion_2D over every (P,R) node),
src/VERB4D_Solver.cpp:851-1216, on the production grid 29x24x101x91
(R x MLT x E x alpha) with synthetic seeded inputs already resident in HBM.
LU caching is off (every step factors all 696 band systems), the plain
FP64 path.

Prints one JSON line. Besides the driver's fields it carries
  roofline      the dominant kernel's achieved rate (algorithmic work per
                launch / its average launch time, HIP events on the library
                stream) against the MI355X peak
                and its HBM traffic per launch, taken from the FETCH_SIZE /
                WRITE_SIZE profile of the same command kept in
                profiles/step_production_grid.md (null when absent)
  cpu_baseline  the reference's own CPU code (oracle/_ref, compiled from the
                PRBEM/VERB-4D sources, OpenMP + OpenBLAS) timed on this
                host on a bounded sample of the same step
"""
import os

os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")  # reference: single-threaded BLAS per call

import argparse  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "verb-4d_amd"))

METRIC = "full-grid steps/s at 29×24×101×91, ≤1e-8 rel err vs FP64; % of roofline"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


PHASE_KERNEL = {"local_factor": "band_lu_window", "local_solve": "band_backsolve"}


def profiled_traffic(phase):
    """HBM bytes per launch of the phase's kernel, as measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/step_production_grid.md
    (calibrated against a copy of known size), or None."""
    kern = PHASE_KERNEL.get(phase)
    path = os.path.join(ROOT, "profiles", "step_production_grid.md")
    if kern is None or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kern:
                return float(parts[2])
    return None


def lu_flops(n, kl, ku):
    """Structural band-LU flops without fill: sum_j (l_j + 2 l_j u_j)."""
    import numpy as np

    j = np.arange(n, dtype=np.float64)
    l = np.minimum(kl, n - 1 - j)
    u = np.minimum(ku, n - 1 - j)
    return float(np.sum(l + 2 * l * u))


def phase_work(w, fused, to_psd=False):
    """Algorithmic work per launch of each kernel phase: (amount, unit, bound).
    fused: the back-substitution runs inside the persistent LU kernel (the
    local_factor phase), and the local_solve phase is the PSD scatter only;
    to_psd: that back-substitution also writes x straight into the PSD, so
    the local_solve phase launches nothing (None)."""
    n = w.nV * w.nK
    kl = ku = w.nV + 1
    nodes = w.nP * w.nR
    pts = w.points
    if fused:
        return {
            "convection": (56.0 * pts, "B", "hbm"),
            "radial": (56.0 * pts, "B", "hbm"),
            "local_assemble": (88.0 * pts, "B", "hbm"),
            # LU flops; the same launch also streams U back for x = U^-1 y
            "local_factor": (lu_flops(n, kl, ku) * nodes, "flop", "mfma"),
            # x read, PSD written
            "local_solve": None if to_psd else (16.0 * pts, "B", "hbm"),
        }
    return {
        # PSD read+write, VP, VR, G_conv, losses, sources: 7 doubles / point
        "convection": (56.0 * pts, "B", "hbm"),
        # PSD r/w, D_LL, G_radial, L grid, sources, losses: 7 doubles / point
        "radial": (56.0 * pts, "B", "hbm"),
        # 8 inputs + 3 outputs per point of the node slices
        "local_assemble": (88.0 * pts, "B", "hbm"),
        "local_factor": (lu_flops(n, kl, ku) * nodes, "flop", "mfma"),
        # U read once: kl+ku+1 doubles per column, plus right-hand side and PSD r/w
        "local_solve": ((kl + ku + 1) * 8.0 * n * nodes + 16.0 * pts, "B", "hbm"),
    }


def host_cpu():
    """(physical cores this process may run on, logical CPUs it may run on,
    CPUs of the machine, CPU model)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            pkg = open(base + "physical_package_id").read().strip()
            core = open(base + "core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return len(cores), len(cpus), os.cpu_count(), model


HOST_CPU = None  # host_cpu() taken before anything else pins this process's threads


def cpu_threads(requested):
    """The baseline's OpenMP threads: the physical cores this process may
    use, within OMP_NUM_THREADS when the host sets a share."""
    phys, logical, _, _ = HOST_CPU
    n = phys
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    if requested > 0:
        n = min(n, requested)
    return max(1, n)


def node_errors(got, ref, floor=1e-12):
    """(largest per-(P,R)-node normwise error, largest pointwise relative
    error over points with |ref| >= floor * the node's max|ref|)."""
    import numpy as np

    num = np.abs(got - ref)
    den = np.abs(ref).max(axis=(2, 3), keepdims=True)
    den = np.where(den == 0, 1.0, den)
    normwise = float((num.max(axis=(2, 3), keepdims=True) / den).max())
    mask = np.abs(ref) >= floor * den
    point = float((num[mask] / np.abs(ref[mask])).max()) if mask.any() else 0.0
    return normwise, point


def cpu_baseline(w, psd, threads, gpu_step, cached_nodes=64):
    """The reference CPU path (oracle/_ref, the PRBEM/VERB-4D sources built
    with OpenMP and OpenBLAS) on one full step of the timed workload from the
    state psd: convection on every (V,K) plane, radial diffusion on every
    line, local diffusion on every node, each operator timed. Its output is
    also the check of the GPU step from the same state (gpu_step(psd) ->
    PSD). With LU caching on, the local step's cached-factor solves are timed
    on cached_nodes nodes (their factors computed first, untimed) and scaled
    to the grid."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if not O.ref_available():
        return None, None
    # The reference runs in a child process of its own: its OpenMP runtime
    # starts there with OMP_PROC_BIND=close and the thread count, and no
    # thread placement of this (GPU) process constrains it.
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.npy"), os.path.join(tmp, "out.npy")
        np.save(fin, psd)
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS="1")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", fin, fout, str(threads)],
                           env=env, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            return None, None
        t = json.loads(r.stdout.strip().splitlines()[-1])
        x = np.load(fout)
    est = t["conv"] + t["radial"] + t["local"]
    got = gpu_step(psd)
    normwise, point = node_errors(got, x)
    check = {"max_node_err": normwise, "max_pointwise_rel_err": point, "pointwise_floor": 1e-12,
             "against": "reference CPU step (OpenBLAS dgbtrf/dgbtrs) from the same state",
             "bar": 1e-8}
    cached = None
    if t.get("cached_local") is not None:
        tl = t["cached_local"]
        cached = {"value": 1.0 / (t["conv"] + t["radial"] + tl), "unit": "steps/s",
                  "seconds_per_step": t["conv"] + t["radial"] + tl,
                  "sample": "reference LU_CACHING build: convection and radial as above, cached-factor solves of "
                            "%d of %d nodes in %.2f s scaled to the grid" % (t["cached_nodes"], w.nP * w.nR,
                                                                              t["cached_sample_s"])}
    phys, logical, ncpu, model = HOST_CPU
    base = {"value": 1.0 / est, "unit": "steps/s", "cores": threads, "kind": "reference",
            "sample": "reference CPU path (OpenMP schedule(dynamic,1), OMP_PROC_BIND=close, OpenBLAS dgbtrf/dgbtrs "
                      "single-threaded per call), LU caching off: one full step, convection %.2f s (all %d planes), "
                      "radial %.2f s (all %d lines), local %.2f s (all %d nodes)"
                      % (t["conv"], w.nV * w.nK, t["radial"], w.nP * w.nV * w.nK, t["local"], w.nP * w.nR),
            "seconds_per_step": est, "host_cpu": model, "physical_cores": phys, "logical_cpus": logical,
            "nproc": ncpu, "lu_caching_on": cached}
    return base, check


def cpu_child(fin, fout, threads, cached_nodes=64):
    """The reference step of cpu_baseline, in its own process (no GPU):
    one full step of the production workload from the state in fin, each
    operator timed; the result goes to fout, the times to stdout (JSON)."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from verb4d import workload as W

    w = W.config1(seed=1001)
    psd = np.load(fin)
    x = psd.copy()
    t = {}
    for op, args in (("conv", (-1, 0, 0)), ("radial", (0, -1, 0)), ("local", (0, 0, -1))):
        t0 = time.perf_counter()
        O.ref_step_sample(w, x, threads, *args)
        t[op] = time.perf_counter() - t0
    np.save(fout, x)
    t["cached_local"] = None
    if O.ref_available("lucache"):
        nn = min(cached_nodes, w.nP * w.nR)
        n = w.nV * w.nK
        ldab = 3 * (w.nV + 1) + 1
        lu = np.zeros(ldab * n * nn)
        piv = np.zeros(n * nn, dtype=np.int64)
        h = O.ref_lib("lucache")
        g = O.grid_struct(w)
        y = psd.copy()
        h.ref_step_local_cached(ctypes.byref(g), O._p(y), threads, O._p(lu), piv.ctypes.data_as(O._i64p), 1, 0, nn)
        t0 = time.perf_counter()
        h.ref_step_local_cached(ctypes.byref(g), O._p(y), threads, O._p(lu), piv.ctypes.data_as(O._i64p), 0, 0, nn)
        ts = time.perf_counter() - t0
        t["cached_local"] = ts * (w.nP * w.nR) / nn
        t["cached_nodes"] = nn
        t["cached_sample_s"] = ts
    print(json.dumps(t), flush=True)


def mixed_check(w3, psd_in, got, threads, nodes=16):
    """The reference's Diffusion_2D (oracle/_ref, FP64 OpenBLAS dgbtrf) on
    `nodes` nodes of the stiff grid spread over (P,R), in parallel, against
    the mixed-precision local solve's PSD from the same input."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if not O.ref_available():
        return None
    total = w3.nP * w3.nR
    pick = [int(i) for i in np.linspace(0, total - 1, nodes)]

    def one(q):
        iP, iR = divmod(q, w3.nR)
        o = (iP, iR)
        ref = O.diffusion_2d(psd_in[o], w3.f["Vg"][o], w3.f["Kg"][o], w3.nV, w3.nK, w3.bc["Vl"][o], w3.bc["Vu"][o],
                             w3.bc["Kl"][o], w3.bc["Ku"][o], (1, 1, 1, 2), w3.f["DVV"][o], w3.f["DKK"][o],
                             w3.f["DVK"][o], w3.f["DVK"][o], w3.f["G_local"][o], w3.f["sources"][o],
                             w3.f["losses_local"][o], w3.dt, w3.sub_dt, lib="ref")
        return node_errors(got[iP:iP + 1, iR:iR + 1], ref[None, None])

    with ThreadPoolExecutor(max(1, threads)) as ex:
        errs = list(ex.map(one, pick))
    return {"max_node_err": max(e[0] for e in errs), "max_pointwise_rel_err": max(e[1] for e in errs),
            "pointwise_floor": 1e-12, "nodes_checked": len(pick),
            "against": "reference Diffusion_2D (FP64, OpenBLAS dgbtrf/dgbtrs) on nodes spread over (P,R)"}


def learned_sequence(ctx, ext, ev0, ev1, nsteps, tsteps):
    """The learned warm start on the held-out sequence: the production grid
    (seed 4001, not in the training data) with Kp 1, 8 and 2 for a third of
    the steps each. Two copies of the PSD advance side by side: one by the
    learned step (network x0, FP32 GMRES judged by the FP64 residual at 1e-10,
    band LU for every node not converged within 8 inner iterations), one by
    the plain FP64 step (the parity-checked band LU path). Reports the
    fallback rate per Kp segment, the largest per-node normwise and
    pointwise (|ref| >= 1e-12 node max) difference to the FP64 path over the
    sequence, and steps/s of both paths timed on tsteps steps at each Kp."""
    import torch

    from verb4d import learned as L
    from verb4d import workload as W
    from verb4d.device import DeviceGrid

    if not os.path.exists(L.WEIGHTS):
        return None
    params, meta = L.load()
    w = W.config1(seed=4001)
    w.set_kp(1.0)
    dl = DeviceGrid(w, ctx=ctx)
    dp = DeviceGrid(w, ctx=ctx)
    net = L.UNet(ctx, 696, 104, 96, train=False, params=params)
    prm = DeviceGrid.gmres_params_capped()
    fields = ["DKK", "DVV", "DVK", "DLL", "VP", "VR"]
    seg = {}
    worst_n, worst_p = 0.0, 0.0
    third = max(1, nsteps // 3)
    kps = [1.0, 8.0, 2.0]
    for it in range(nsteps):
        kp = kps[min(2, it // third)]
        if kp != w.kp:
            w.set_kp(kp)
            dl.refresh(fields)
            dp.refresh(fields)
        st = L.step_learned(dl, net, prm, stats=True)
        dp.step()
        s = seg.setdefault("kp%g" % kp, {"steps": 0, "nodes": 0, "fallback": 0, "iterations": 0})
        s["steps"] += 1
        s["nodes"] += st["nodes"]
        s["fallback"] += st["fallback"]
        s["iterations"] += st["iterations"]
        if it % 16 == 15 or it == nsteps - 1:
            a, b = dl.psd_host(), dp.psd_host()
            en, ep = node_errors(a, b)
            worst_n, worst_p = max(worst_n, en), max(worst_p, ep)
    dl.check_status()
    dp.check_status()
    for s in seg.values():
        s["fallback_rate"] = round(s["fallback"] / max(1, s["nodes"]), 4)
        s["mean_iterations"] = round(s["iterations"] / max(1, s["nodes"]), 2)
    # timing at each Kp, without per-step statistics
    timing = {}
    for kp in kps:
        w.set_kp(kp)
        dl.refresh(fields)
        dp.refresh(fields)
        for name, fn in (("learned", lambda: L.step_learned(dl, net, prm)), ("plain", lambda: dp.step())):
            fn()
            ctx.synchronize()
            ev0.record(ext)
            for _ in range(tsteps):
                fn()
            ev1.record(ext)
            ctx.synchronize()
            torch.cuda.synchronize()
            timing.setdefault(name, {})["kp%g" % kp] = round(1000.0 * tsteps / ev0.elapsed_time(ev1), 4)
    phases = {}
    ctx.set_timing(True)
    ctx.timings()
    L.step_learned(dl, net, prm)
    for k, (ms, cnt) in ctx.timings().items():
        if cnt:
            phases[k] = round(ms, 3)
    ctx.set_timing(False)
    mean = lambda d: sum(d.values()) / len(d)  # noqa: E731
    return {"workload": "production grid 29x24x101x91, seed 4001, Kp 1 -> 8 -> 2 over %d steps" % nsteps,
            "rule": "x0 from the U-Net, FP32 GMRES (2 cycles of 4) judged by the FP64 residual at 1e-10, band LU "
                    "for nodes not converged within 8 inner iterations",
            "segments": seg, "max_node_err_vs_fp64": worst_n, "max_pointwise_rel_err_vs_fp64": worst_p,
            "steps_per_s": timing["learned"], "plain_steps_per_s": timing["plain"],
            "speedup_vs_plain": round(mean(timing["learned"]) / mean(timing["plain"]), 3),
            "phase_ms_one_step": phases,
            "network": {"parameters": int(params.size), "training_samples": int(meta.get("samples", 0)),
                        "epochs": int(meta.get("epochs", 0))}}


def copy_calibration():
    import torch
    n = 64 << 20
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy calibration: %d bytes read and %d bytes written per copy launch" % (n * 8, n * 8), flush=True)


def main():
    global HOST_CPU
    HOST_CPU = host_cpu()
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-child":
        cpu_child(sys.argv[2], sys.argv[3], int(sys.argv[4]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: the physical cores available, within OMP_NUM_THREADS)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the stiff 29x24x151x121 grid solved by FP32 LU + FP64 refinement")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned warm start on the held-out Kp 1 -> 8 -> 2 sequence")
    ap.add_argument("--learned-steps", type=int, default=288,
                    help="steps of the held-out Kp sequence (thirds at Kp 1, 8, 2)")
    ap.add_argument("--no-cached", action="store_true",
                    help="skip timing steps that reuse cached LU factors (the reference's LU_CACHING mode)")
    ap.add_argument("--no-remap", action="store_true", help="skip the adiabatic remapping line")
    ap.add_argument("--no-da", action="store_true", help="skip the data assimilation line")
    ap.add_argument("--no-gmres", action="store_true",
                    help="skip timing steps whose local solve is the matrix-free GMRES path")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shapes TQ,RB,CB,PERSISTENT instead of the autotuner (profiling runs)")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-extent production grid (loss-cone pitch-angle extents, 4040..9191 unknowns)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eagerly launched steps instead of the captured HIP graph of the step")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy three times (a known-size kernel against "
                         "which the FETCH_SIZE / WRITE_SIZE counters are scaled) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported (the multi-GPU "
              "sharding is a design note, not an implemented path)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    import verb4d
    from verb4d import workload as W
    from verb4d.device import DeviceGrid

    torch.cuda.set_device(0)
    w = W.config1(seed=1001)  # Kp = 2 regime of the production configuration
    ctx = verb4d.Context()
    dg = DeviceGrid(w, ctx=ctx)
    print("bench: workload ready (%d points, %d nodes)" % (w.points, w.nP * w.nR), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        dg.step()
    dg.check_status()
    torch.cuda.synchronize()
    ctx.synchronize()
    tuning = None
    if args.tuning:
        tq, rb, cb, pers = (int(v) for v in args.tuning.split(","))
        dg.set_tuning(convection_tq=tq, lu_rb=rb, lu_cb=cb, lu_persistent=pers)
        tuning = {"convection_tq": tq, "lu_rb": rb, "lu_cb": cb, "lu_persistent": pers, "fixed": True}
    elif not args.no_autotune:
        tuning = dg.autotune()  # launch shapes only; psd is restored
        print("bench: autotuned %s" % tuning, file=sys.stderr, flush=True)

    ext = torch.cuda.ExternalStream(ctx.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ctx.set_timing(True)
    ctx.timings()  # clear
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    ev0.record(ext)
    for _ in range(args.steps):
        dg.step()
    ev1.record(ext)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    phases = ctx.timings()
    ctx.set_timing(False)
    dg.check_status()
    ms_events = ev0.elapsed_time(ev1)
    ms_per_step = ms_events / args.steps
    print("bench: %d eager steps with phase events, %.3f ms/step, wall %.3f s" % (args.steps, ms_per_step, wall),
          file=sys.stderr, flush=True)
    eager_ms_per_step = ms_per_step
    if not args.no_graph:
        # the same step replayed from its captured HIP graph: the headline
        dg.step_graph()  # eager step + capture
        for _ in range(max(1, args.warmup)):
            dg.step_graph()
        dg.check_status()
        ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(ext)
        for _ in range(args.steps):
            dg.step_graph()
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        dg.check_status()
        ms_per_step = ev0.elapsed_time(ev1) / args.steps
        print("bench: %d graph-replayed steps, %.3f ms/step, wall %.3f s" % (args.steps, ms_per_step, wall),
              file=sys.stderr, flush=True)

    # the autotuner's persistent LU queue runs the back-substitution in the
    # same launch
    fused = bool((tuning or {}).get("lu_persistent", 0))
    # register windows of 4 or more wavefronts hold the back-substitution,
    # which then writes x into the PSD itself
    to_psd = fused and (tuning or {}).get("lu_rb", 0) >= 4
    work = phase_work(w, fused, to_psd)
    phase_out = {}
    dom, dom_ms = None, -1.0
    for p, (ms, cnt) in phases.items():
        if cnt == 0 or work[p] is None:
            continue
        avg = ms / cnt
        amount, unit, bound = work[p]
        rate = amount / (avg * 1e-3)
        if unit == "flop":
            ach, peak, u = rate / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
        else:
            ach, peak, u = rate / 1e9, PEAK_HBM_GBS, "GB/s"
        phase_out[p] = {"avg_ms": round(avg, 4), "achieved": round(ach, 3), "unit": u, "frac": round(ach / peak, 4)}
        if avg > dom_ms:
            dom, dom_ms = p, avg
    amount, unit, bound = work[dom]
    d = phase_out[dom]
    roofline = {"bound": bound, "achieved": d["achieved"],
                "peak": PEAK_FP64_TFLOPS if unit == "flop" else PEAK_HBM_GBS, "unit": d["unit"],
                "frac": d["frac"], "traffic": profiled_traffic(dom), "kernel": dom}
    if fused and dom == "local_factor":
        roofline["note"] = ("band_lu_window with the back-substitution fused in (the persistent launch solves "
                            "x = U^-1 y on the workgroups that run out of nodes to factor%s); achieved = LU flops / "
                            "launch time" % (", writing x into the PSD" if to_psd else ""))

    cached = None
    if not args.no_cached:
        # LU caching on: factors from the first step are reused (the
        # coefficients do not change between Kp updates)
        dg.set_psd(w.psd0)
        dg.alloc_lu_cache()
        dg.step(use_cache=True, recompute_lu=True)
        ctx.synchronize()
        ev0.record(ext)
        for _ in range(args.steps):
            dg.step(use_cache=True, recompute_lu=False)
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        cached = {"steps_per_s": round(1000.0 * args.steps / ev0.elapsed_time(ev1), 4),
                  "ms_per_step": round(ev0.elapsed_time(ev1) / args.steps, 4)}
        dg.lu_cache = None
        dg.ipiv_cache = None
        torch.cuda.empty_cache()

    remap = None
    if not args.no_remap:
        # adiabatic remapping of the PSD between two L* fields (compression
        # 0.05 <-> 0.12); per point it reads L*_old, L*_new and the PSD and
        # writes the PSD: 32 B
        lo = torch.from_numpy(W.lstar(w, 0.05)).cuda()
        ln = torch.from_numpy(W.lstar(w, 0.12)).cuda()
        dg.set_psd(w.psd0)
        dg.remap(lo, ln)
        dg.check_status()
        reps = max(10, args.steps)
        ev0.record(ext)
        for i in range(reps):
            a, b = (lo, ln) if i % 2 == 0 else (ln, lo)
            ctx.check(verb4d.lib().verb_remap_batched(ctx.handle, dg.psd.data_ptr(), a.data_ptr(), b.data_ptr(), w.nP,
                                                      w.nR, w.nV, w.nK))
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        dg.check_status()
        ms = ev0.elapsed_time(ev1) / reps
        gbs = 32.0 * w.points / (ms * 1e-3) / 1e9
        remap = {"workload": "production grid 29x24x101x91, L* compression 0.05 <-> 0.12 (220584 lines of 29)",
                 "ms": round(ms, 4), "achieved": round(gbs, 2), "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                 "bytes_per_point": 32}
        del lo, ln

    da = None
    if not args.no_da:
        # the Kalman update along L on every (V,K) line of one MLT slice: per
        # line (n = 29 points, m observed) F = B A^-1 (LU of A^T, n solves),
        # F Pa F^T, the gain (LU of S^T, n solves) and K H Pf
        rng = np.random.default_rng(5)
        n = w.nR
        obs = np.full(w.shape, np.nan)
        keep = rng.random(w.shape[1:]) < 0.4
        obs[0] = np.where(keep, w.psd0[0] * (1 + 0.2 * rng.standard_normal(w.shape[1:])), np.nan)
        m = keep.sum(axis=0)  # observed points per line
        tobs = torch.from_numpy(obs).cuda()
        cov = torch.zeros((w.nV, w.nK, n, n), dtype=torch.float64, device="cuda")
        sc = float(np.abs(w.psd0[0]).max())
        dg.set_psd(w.psd0)
        dg.da_radial(tobs, cov, 3600.0, 1e-2 * sc ** 2, 4e-2 * sc ** 2)
        dg.check_status()
        reps = max(10, args.steps)
        from verb4d import DaParams
        prm = DaParams(3600.0, 1e-2 * sc ** 2, 4e-2 * sc ** 2, 0)
        ev0.record(ext)
        for _ in range(reps):
            ctx.check(verb4d.lib().verb_da_radial(ctx.handle, ctypes.byref(dg.grid), dg.psd.data_ptr(), tobs.data_ptr(),
                                                  cov.data_ptr(), None, ctypes.byref(prm)))
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        dg.check_status()
        ms = ev0.elapsed_time(ev1) / reps
        flops = float(np.sum(2 * n ** 3 / 3 + 2 * n ** 3 + 2 * n ** 3 + 2 * n ** 3 + 2 * m ** 3 / 3 + 2 * m ** 2 * n
                             + 2 * n * n * m + 2 * n * m))
        da = {"workload": "production grid, MLT slice 0: 9191 (E, alpha) lines of 29 points, 40 % observed",
              "ms": round(ms, 4), "gflop": round(flops / 1e9, 3), "achieved": round(flops / (ms * 1e-3) / 1e12, 3),
              "unit": "TFLOP/s", "frac": round(flops / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)}
        del tobs, cov

        # the (MLT, R) plane filter on every (V,K) plane: hourly assimilation
        # (dt = 1/24 day, the reference's default step) with the velocities
        # per day, 40 % of the points observed, the loss array the solver
        # passes (its Sources: zero here). Per plane (N = 696, m observed):
        # the model matrix M^nt and M1d = M^3 (GEMMs by repeated squaring),
        # Q = M1d T M1d^T, Pf = M Pa M^T + Q, the gain (LU of S, N solves)
        # and K H Pf; the call that rebuilds M and Q, then one that reuses
        # them (Q kept in its array)
        torch.cuda.empty_cache()
        N, NQ = w.nP * w.nR, w.nV * w.nK
        vp0, vr0 = w.f["VP"], w.f["VR"]
        w.f["VP"], w.f["VR"] = vp0 * 86400.0, vr0 * 86400.0
        dg.refresh(["VP", "VR"])
        dg.set_psd(w.psd0)
        rng = np.random.default_rng(6)
        keep = rng.random(w.shape) < 0.4
        obs = np.where(keep, np.abs(w.psd0 * (1 + 0.2 * rng.standard_normal(w.shape))), np.nan)
        tobs = torch.from_numpy(obs).cuda()
        tloss = torch.from_numpy(np.ascontiguousarray(w.f["sources"])).cuda()
        del obs
        cov = torch.zeros((NQ, N, N), dtype=torch.float64, device="cuda")
        model = torch.empty((NQ, N, N), dtype=torch.float64, device="cuda")
        qn = torch.empty((NQ, N, N), dtype=torch.float64, device="cuda")
        dtd, corr = 1.0 / 24, 2.0
        times = {}
        for name, rec in (("rebuild", True), ("reuse", False)):
            ctx.synchronize()
            torch.cuda.synchronize()
            ev0.record(ext)
            dg.da_convection(tobs, tloss, cov, model, dtd, corr, 1e-2, 4e-2, recompute_model=rec, process_noise=qn)
            ev1.record(ext)
            ctx.synchronize()
            torch.cuda.synchronize()
            dg.check_status()
            times[name] = ev0.elapsed_time(ev1)
        # algorithmic FLOPs (the products of the reference's operations)
        dP = w.f["Pg"][1, 0, 0, 0] - w.f["Pg"][0, 0, 0, 0]
        dR = w.f["Rg"][0, 1, 0, 0] - w.f["Rg"][0, 0, 0, 0]
        cour = (np.abs(w.f["VP"]) * dtd / dP + np.abs(w.f["VR"]) * dtd / dR).max(axis=(0, 1)).ravel()
        nt = np.maximum(1, np.ceil(cour)).astype(int)
        m = keep.sum(axis=(0, 1)).ravel().astype(np.float64)

        def n_gemm(p):
            return 0 if p == 1 else (p.bit_length() - 1) + (bin(p >> 1).count("1") - 1) + (p & 1)

        p1 = int(np.ceil(3.0 / 24 / dtd))
        g = np.array([n_gemm(int(t)) for t in nt]) + n_gemm(p1)
        gemm = 2.0 * N ** 3
        tail = 2 * N * N * m + 2 * m ** 3 / 3 + 2 * m * m * N
        fl = {"rebuild": float(np.sum((g + 3) * gemm + tail)), "reuse": float(np.sum(2 * gemm + tail))}
        da_plane = {"workload": "production grid: 9191 (E, alpha) planes of 24 x 29 = 696 points, 40 %% observed, "
                                "dt = 1/24 day, sub-steps %d-%d, M1d = M^%d" % (nt.min(), nt.max(), p1),
                    "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS}
        for k in ("rebuild", "reuse"):
            ach = fl[k] / (times[k] * 1e-3) / 1e12
            da_plane[k] = {"ms": round(times[k], 3), "tflop": round(fl[k] / 1e12, 3), "achieved": round(ach, 3),
                           "frac": round(ach / PEAK_FP64_TFLOPS, 4)}
        w.f["VP"], w.f["VR"] = vp0, vr0
        dg.refresh(["VP", "VR"])
        dg.set_psd(w.psd0)
        del tobs, tloss, cov, model, qn
        torch.cuda.empty_cache()
        da = {"radial": da, "plane": da_plane}

    gm = None
    gm8 = None
    if not args.no_gmres:
        # local solve by FP32 GMRES cycles with the FP64 residual test
        # (max|b - Ax| <= 1e-11 max|b|) and band-LU fallback
        def time_gmres(prm):
            dg.set_psd(w.psd0)
            st = dg.step_gmres(prm, stats=True)
            ctx.synchronize()
            ev0.record(ext)
            for _ in range(args.steps):
                dg.step_gmres(prm)
            ev1.record(ext)
            ctx.synchronize()
            torch.cuda.synchronize()
            return st, ev0.elapsed_time(ev1) / args.steps

        prm = DeviceGrid.gmres_params(tol=1e-11, restart=20, max_cycles=8)
        st, ms = time_gmres(prm)
        gm = {"steps_per_s": round(1000.0 / ms, 4), "ms_per_step": round(ms, 4), "tol": 1e-11, "first_step": st}
        # the warm-start rule: at most 8 inner iterations, then the band LU
        gm8 = {}
        prm8 = DeviceGrid.gmres_params_capped(tol=1e-10, iterations=8, cycles=2)
        for kp in (2.0, 6.0):
            if kp != w.kp:
                w.set_kp(kp)
                dg.refresh(["DKK", "DVV", "DVK", "DLL", "VP", "VR"])
            st, ms = time_gmres(prm8)
            gm8["kp%g" % kp] = {"steps_per_s": round(1000.0 / ms, 4), "ms_per_step": round(ms, 4),
                                "fallback_rate": round(st["fallback"] / max(1, st["nodes"]), 4), "first_step": st}
        gm8["rule"] = "tol 1e-10 (FP64 residual / max|b|), 8 inner iterations in 2 FP32 cycles, then band LU"
        if w.kp != 2.0:
            w.set_kp(2.0)
            dg.refresh(["DKK", "DVV", "DVK", "DLL", "VP", "VR"])

    uneven = None
    if not args.no_uneven:
        # same grid, pitch-angle extents following the loss cone plus a drift
        # band (4040..9191 unknowns per node), Kp 2 / 6 across MLT; persistent
        # LU queue over nodes sorted by size, replayed from its graph
        w2, k0 = W.config2(seed=2001)
        dg2 = DeviceGrid(w2, ctx=ctx)
        dg2.step_graph(k0)  # eager step + capture
        for _ in range(max(1, args.warmup)):
            dg2.step_graph(k0)
        dg2.check_status()
        ctx.synchronize()
        torch.cuda.synchronize()
        ev0.record(ext)
        for _ in range(args.steps):
            dg2.step_graph(k0)
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        dg2.check_status()
        nn = w2.nV * (w2.nK - k0)
        uneven = {"steps_per_s": round(1000.0 * args.steps / ev0.elapsed_time(ev1), 4),
                  "ms_per_step": round(ev0.elapsed_time(ev1) / args.steps, 4),
                  "unknowns_min": int(nn.min()), "unknowns_max": int(nn.max()),
                  "unknowns_mean": round(float(nn.mean()), 1), "kp": "2 (MLT 0-11), 6 (MLT 12-23)"}
        del dg2
        torch.cuda.empty_cache()

    learned = None
    if not args.no_learned:
        learned = learned_sequence(ctx, ext, ev0, ev1, args.learned_steps, args.steps)

    mixed = None
    if not args.no_mixed:
        # the stiff grid (kl = ku = 152): FP32 band LU + FP64 refinement
        w3 = W.config3(seed=3001)
        dg3 = DeviceGrid(w3, ctx=ctx)
        mst = dg3.local_mixed(tol=1e-12, max_sweeps=8)
        ctx.synchronize()
        psd3_in = w3.psd0
        mixed_out = dg3.psd_host()
        ev0.record(ext)
        for _ in range(args.steps):
            dg3.step_mixed(tol=1e-12, max_sweeps=8)
        ev1.record(ext)
        ctx.synchronize()
        torch.cuda.synchronize()
        ms3 = ev0.elapsed_time(ev1) / args.steps
        mixed = {"workload": "stiff grid 29x24x151x121, 696 band systems of 18271 unknowns (kl=ku=152), dt=600 s",
                 "steps_per_s": round(1000.0 / ms3, 4), "ms_per_step": round(ms3, 4), "tol": 1e-12,
                 "first_local_solve": mst}
        del dg3
        torch.cuda.empty_cache()

    base = None
    check = None
    if not args.no_cpu_baseline:
        threads = cpu_threads(args.cpu_threads)
        psd = dg.psd_host()

        def gpu_step(state):
            dg.set_psd(state)
            dg.step_graph()
            dg.check_status()
            return dg.psd_host()

        base, check = cpu_baseline(w, psd, threads, gpu_step)
        if mixed is not None:
            mixed["check"] = mixed_check(w3, psd3_in, mixed_out, threads)

    value = 1000.0 / ms_per_step
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "production grid 29x24x101x91 (R x MLT x E x alpha), 696 band systems of 9191 "
                               "unknowns (kl=ku=102), Kp=2, dt=300 s, LU caching off",
                   "grid": [w.nR, w.nP, w.nV, w.nK], "lu_caching": False,
                   "launch": "hip graph replay" if not args.no_graph else "eager",
                   "tuning": tuning},
        "roofline": roofline,
        "cpu_baseline": base,
        "phases": phase_out,
        "eager_ms_per_step": round(eager_ms_per_step, 4),
    }
    if check is not None:
        out["validation"] = check
    if cached is not None:
        out["lu_cached"] = cached
    if gm is not None:
        out["gmres_local"] = gm
    if gm8 is not None:
        out["gmres_8_iterations"] = gm8
    if mixed is not None:
        out["mixed_precision_stiff"] = mixed
    if learned is not None:
        out["learned"] = learned
    if uneven is not None:
        out["uneven_extents"] = uneven
    if remap is not None:
        out["remap"] = remap
    if da is not None:
        out["data_assimilation"] = da
    if base:
        out["speedup_vs_cpu"] = round(value / base["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
