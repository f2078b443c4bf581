"""Train the warm-start network on pairs the GPU solver generates, and save its weights.

    python -m nextsimdg_amd.train_warm_start [--nx 256] [--tracks 200] [--per-track 10]
        [--epochs 30] [--out nextsimdg_amd/data/warm_start_256.npy]

Pairs: for each seeded cyclone track (ice cycling through plain, heterogeneous and
leads) the plain model runs; at sampled steps the start state is the input and the
velocity mEVP reaches after up to 10^4 subcycles (or r <= 0.05) the target. The last
tenth of the tracks is held out. Training: Adam (lr 1e-3), batch 16, BF16 MFMA kernels
of libnsdg_hip.so. Prints one progress line per track and per epoch and writes the
weights (FP32, flat order of include/nsdg_cnn.h) with a JSON summary beside them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--tracks", type=int, default=200)
    ap.add_argument("--per-track", type=int, default=10)
    ap.add_argument("--converge", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=os.path.join(HERE, "data", "warm_start_256.npy"))
    args = ap.parse_args()

    from nextsimdg_amd import learned as L

    def log(m):
        print(m, flush=True)

    t0 = time.time()
    X, T, W, stats = L.generate_pairs(args.nx, list(range(args.tracks)), args.per_track, converge=args.converge,
                                      r_stop=0.05, check_every=2000, log=log)
    tgen = time.time() - t0
    n = X.shape[0]
    ntest = max(args.per_track, (args.tracks // 10) * args.per_track)
    train_idx, test_idx = list(range(n - ntest)), list(range(n - ntest, n))
    net = L.WarmStartNet(args.seed)
    t0 = time.time()
    hist = L.train(net, X, T, W, args.epochs, args.batch, args.lr, seed=args.seed, idx=train_idx, log=log)
    ttrain = time.time() - t0
    ln, lz = L.evaluate(net, X, T, W, test_idx)
    log(f"held-out loss: network {ln:.4e}, cold start {lz:.4e} (ratio {ln / lz:.4f})")
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, net.get_params())
    summary = {
        "grid": args.nx, "tracks": args.tracks, "pairs": n, "held_out_pairs": ntest, "epochs": args.epochs,
        "batch": args.batch, "lr": args.lr, "generation_s": round(tgen, 1), "training_s": round(ttrain, 1),
        "loss_history": [float(f"{v:.5e}") for v in hist],
        "held_out_loss_network": ln, "held_out_loss_cold_start": lz, "held_out_ratio": ln / lz,
        "residual_of_targets": {"median_r_start": float(np.median([s[3] for s in stats])),
                                "median_r_target": float(np.median([s[5] for s in stats]))},
    }
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps({k: v for k, v in summary.items() if k != "loss_history"}))


if __name__ == "__main__":
    main()
