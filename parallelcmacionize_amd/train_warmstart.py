"""Train the warm-start network on fields the method converges itself.

    python -m parallelcmacionize_amd.train_warmstart --problems 32 \\
        --out parallelcmacionize_amd/data/warmstart_clumpy256.npz

Training problems are 256^3 clumpy media like the 256^3 clumpy workload
(same box, statistics and source ranges; field and source seeds 1000+,
2000+, 3000+ so the workload itself is never seen), each converged by the
plain iteration (1e8 packets per iteration, mean |dx_n| < 1e-3) and cut into
64 crops of 64^3.  The last problem's crops are held out for validation.
Training: Adam on random 32^3 crops, BF16 MFMA kernels with FP32
accumulation, all on the GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import warmstart as WS


def log(*a):
    print(*a, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=32)
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--ncell", type=int, default=256)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--train-crop", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", required=True)
    ap.add_argument("--offsets", type=float, nargs="*",
                    default=[0.0, -1.0, -2.0, -3.0, -4.0])
    ap.add_argument("--save-data", default=None,
                    help="also write the training crops to this .npz")
    a = ap.parse_args(argv)

    t0 = time.time()
    feats, tgts, recs = WS.make_training_set(range(a.problems), int(a.packets),
                                             crop=a.crop, ncell=a.ncell,
                                             log=log)
    log("training set: %d crops of %d^3 in %.1f s" % (len(feats), a.crop,
                                                      time.time() - t0))
    if a.save_data:
        np.savez(a.save_data, features=feats, targets=tgts)
    per = (a.ncell // a.crop) ** 3
    ntrain = len(feats) - per if a.problems > 1 else len(feats)
    f_tr, t_tr = feats[:ntrain], tgts[:ntrain]
    f_va, t_va = feats[ntrain:], tgts[ntrain:]
    # FP64 statistics: a float32 reduction over 1e8 values drifts badly
    mean = f_tr.reshape(-1, 2).mean(axis=0, dtype=np.float64)
    std = f_tr.reshape(-1, 2).std(axis=0, dtype=np.float64)
    net = WS.new_net(a.seed, mean, 1.0 / std)
    for i in range(0, ntrain, 64):
        net.add_samples(f_tr[i:i + 64], t_tr[i:i + 64])
    t1 = time.time()
    crop = [a.train_crop] * 3
    phases = [(0.6, a.lr), (0.3, a.lr * 0.3), (0.1, a.lr * 0.1)]
    losses = []
    for frac, lr in phases:
        n = max(1, int(round(frac * a.steps)))
        l = net.train(n, a.batch, crop, lr, seed=a.seed + len(losses))
        losses.append(l)
        log("trained %d steps at lr %.1e: loss %.4f -> %.4f"
            % (n, lr, l[:20].mean(), l[-20:].mean()))
    losses = np.concatenate(losses)
    t_train = time.time() - t1

    def val_loss():
        if len(f_va) == 0:
            return None
        v = [net.loss_and_grad(f_va[i:i + 1], t_va[i:i + 1])[0]
             for i in range(len(f_va))]
        return float(np.mean(v))

    vl = val_loss()
    log("validation loss %s" % vl)

    # the logit offset: iterations to convergence on a held-out problem for
    # a few shifts toward ionized (a front predicted too close to its source
    # costs about one iteration per cell, one predicted too far converges
    # geometrically), the best one baked into the last bias
    box, n, pos, L = WS.clumpy_problem(a.ncell, field_seed=999, src_seed=998,
                                       lum_seed=997)
    prob = WS.Problem(box, n, pos, L)
    plain = WS.converge_plain(prob, int(a.packets), seed=31)
    log("held-out problem, plain: %d iterations %.3f s"
        % (plain["iterations"], plain["total_seconds"]))
    trials = []
    applied = 0.0
    for delta in a.offsets:
        WS.shift_logits(net, delta - applied)
        applied = delta
        w = WS.converge_warm(prob, net, int(a.packets), seed=31)
        trials.append({"offset": delta, "iterations": w["iterations"],
                       "seconds": w["total_seconds"], "l1": w["l1"]})
        log("offset %+.1f: %d iterations %.3f s" % (delta, w["iterations"],
                                                   w["total_seconds"]))
    best = min(trials, key=lambda t: (t["iterations"], t["seconds"]))
    WS.shift_logits(net, best["offset"] - applied)
    prob.close()
    meta = {"problems": a.problems, "packets": a.packets, "ncell": a.ncell,
            "crop": a.crop, "crops": int(len(feats)), "train_crops": int(ntrain),
            "steps": a.steps, "batch": a.batch, "train_crop": a.train_crop,
            "lr": a.lr, "train_seconds": t_train,
            "final_train_loss": float(losses[-50:].mean()),
            "validation_loss": vl, "loss_units": "(logit / 8)^2",
            "logit_offset": best["offset"], "offset_trials": trials,
            "heldout_plain_iterations": plain["iterations"],
            "heldout_plain_seconds": plain["total_seconds"],
            "problem_records": recs}
    log(json.dumps({k: v for k, v in meta.items() if k != "problem_records"}))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    WS.save_net(net, a.out, meta)
    net.close()


if __name__ == "__main__":
    sys.exit(main())
