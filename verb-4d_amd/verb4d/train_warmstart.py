"""Trains the warm-start U-Net on the GPU with the library's own kernels.

Training pairs come from the plain solver: for each workload below, every
step runs convection and radial diffusion, records each (P,R) node's
features (log Daa, log DEE, DaE/sqrt(Daa DEE), log PSD_in, dt) and
PSD_in / s, runs the local solve and records PSD_out / s (s = the node's
max |PSD_in|). Workloads (fresh seeds, none used by the tests or the bench):
  production grid 29x24x101x91 at Kp 2 and Kp 6 (the plain FP64 band LU),
  the uneven-extent grid (Kp 2 / 6 across MLT; the pitch-angle columns a
  node leaves untouched are masked out of the loss),
  the stiff grid 29x24x151x121 (FP32 LU + FP64 refinement to 1e-12).
About 20k node-steps in all. Adam with a cosine learning-rate decay; 5% of
the samples are held out and their loss is reported per epoch.

    python verb-4d_amd/verb4d/train_warmstart.py --out verb-4d_amd/weights/warmstart_unet.npz
"""
import argparse
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402


def log(*a):
    print(*a, flush=True)


def generate(ctx, net_for, plan):
    """Runs the plain solver along `plan` and returns {(H, W): [feat, pin,
    pout]} device tensors (one group per padded grid size)."""
    import torch

    from verb4d import learned as L
    from verb4d import workload as W
    from verb4d.device import DeviceGrid

    groups = {}
    for name, make, kps, steps in plan:
        t0 = time.time()
        w, k0 = make()
        dg = DeviceGrid(w, ctx=ctx)
        Hp, Wp = L.padded(w.nV, w.nK)
        net = net_for(Hp, Wp)
        nodes = w.nP * w.nR
        for kp in kps:
            if kp is not None and kp != w.kp:
                w.set_kp(kp)
                dg.refresh(["DKK", "DVV", "DVK", "DLL", "VP", "VR"])
            dg.set_psd(w.psd0)
            for s in range(steps):
                dg.step(1, 1, 0)
                feat, pin, scale = net.features(dg, 0, nodes)
                if k0 is not None:
                    dg.local_uneven(k0)
                elif w.nV + 1 > 112:
                    dg.local_mixed(tol=1e-12, max_sweeps=12, stats=False)
                else:
                    dg.local()
                dg.check_status()
                pout = net.normalize(dg, 0, nodes, scale)
                if k0 is not None:
                    # columns below k0 are not part of the node's system
                    mask = torch.ones((nodes, Hp, Wp), dtype=torch.float32, device="cuda")
                    kk = torch.from_numpy(np.asarray(k0, dtype=np.int64).ravel()).cuda()
                    cols = torch.arange(Wp, device="cuda")[None, None, :]
                    mask = (cols >= kk[:, None, None]).float() * mask
                    pin = pin * mask
                    pout = pout * mask
                g = groups.setdefault((Hp, Wp), [[], [], []])
                g[0].append(feat)
                g[1].append(pin)
                g[2].append(pout)
        log("data: %s kp=%s steps=%d -> %d node-steps in %.1f s" % (name, kps, steps, len(kps) * steps * nodes,
                                                                   time.time() - t0))
        del dg
        torch.cuda.empty_cache()
    return {k: [torch.cat(v[0]), torch.cat(v[1]), torch.cat(v[2])] for k, v in groups.items()}


def default_plan(scale=1.0):
    from verb4d import workload as W

    def c1(seed):
        return lambda: (W.config1(seed=seed), None)

    def c2(seed):
        return lambda: W.config2(seed=seed)

    def c3(seed):
        return lambda: (W.config3(seed=seed), None)

    n = lambda k: max(1, int(round(k * scale)))  # noqa: E731
    return [
        ("production grid", c1(5001), (2.0, 6.0), n(4)),
        ("production grid", c1(5002), (2.0, 6.0), n(3)),
        ("uneven extents", c2(6001), (None,), n(10)),
        ("stiff grid", c3(7001), (None,), n(4)),
    ]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(HERE), "weights", "warmstart_unet.npz"))
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--data-scale", type=float, default=1.0, help="scale the number of solver steps per workload")
    ap.add_argument("--max-seconds", type=float, default=600.0, help="stop training after this long")
    args = ap.parse_args()

    import torch

    import verb4d
    from verb4d import learned as L

    torch.cuda.set_device(0)
    ctx = verb4d.Context()
    nets = {}

    def net_for(H, W):
        if (H, W) not in nets:
            # features and targets only: no forward pass, so one sample of
            # activation storage is enough
            nets[(H, W)] = L.UNet(ctx, 1, H, W, train=False)
        return nets[(H, W)]

    t0 = time.time()
    data = generate(ctx, net_for, default_plan(args.data_scale))
    del nets
    total = sum(v[0].shape[0] for v in data.values())
    log("data: %d node-steps in %.1f s; groups %s" % (total, time.time() - t0,
                                                       {k: v[0].shape[0] for k, v in data.items()}))
    rng = np.random.default_rng(args.seed)
    # hold out 5% of every group
    split = {}
    for k, (f, pi, po) in data.items():
        idx = rng.permutation(f.shape[0])
        nv = max(1, f.shape[0] // 20)
        split[k] = (torch.from_numpy(idx[nv:]).cuda(), torch.from_numpy(idx[:nv]).cuda())
    Hmax = max(k[0] for k in data)
    Wmax = max(k[1] for k in data)
    net = L.UNet(ctx, args.batch, Hmax, Wmax, train=True, params=L.init_params(seed=args.seed))
    steps_per_epoch = sum(len(split[k][0]) // args.batch for k in data)
    total_steps = steps_per_epoch * args.epochs
    step = 0
    history = []
    t_train = time.time()

    def val_loss():
        out, cnt = 0.0, 0
        for k, (f, pi, po) in data.items():
            vi = split[k][1]
            for b in range(0, len(vi) - args.batch + 1, args.batch):
                sel = vi[b:b + args.batch]
                _, ls = net.gradients(f[sel].contiguous(), pi[sel].contiguous(), po[sel].contiguous())
                out += ls
                cnt += 1
        return out / max(1, cnt)

    v0 = val_loss()
    log("epoch 0: held-out loss %.4e" % v0)
    history.append((0, float("nan"), v0))
    stop = False
    for ep in range(1, args.epochs + 1):
        order = []
        for k in data:
            ti = split[k][0][torch.randperm(len(split[k][0]), device="cuda")]
            for b in range(0, len(ti) - args.batch + 1, args.batch):
                order.append((k, ti[b:b + args.batch]))
        perm = rng.permutation(len(order))
        tl, tn = 0.0, 0
        for i in perm:
            k, sel = order[i]
            f, pi, po = data[k]
            lr = args.lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total_steps)))
            want = (tn % 20 == 0)
            ls = net.train_step(f[sel].contiguous(), pi[sel].contiguous(), po[sel].contiguous(), max(lr, 1e-6),
                                want_loss=want)
            if want:
                tl += ls
            tn += 1
            step += 1
            if time.time() - t_train > args.max_seconds:
                stop = True
                break
        vl = val_loss()
        tr = tl / max(1, (tn + 19) // 20)
        history.append((ep, tr, vl))
        log("epoch %d: train loss %.4e held-out loss %.4e (%d steps, %.1f s)" % (ep, tr, vl, step,
                                                                                  time.time() - t_train))
        if stop:
            log("time limit reached")
            break
    params = net.get_params()
    L.save(args.out, params, history=np.array(history), samples=total, epochs=len(history) - 1,
           train_seconds=time.time() - t_train, batch=args.batch, lr=args.lr)
    log("wrote %s (%d parameters)" % (args.out, params.size))


if __name__ == "__main__":
    main()
