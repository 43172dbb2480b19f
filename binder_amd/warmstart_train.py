"""Training of the learned warm start on seeded synthetic 160x192x160 pairs.

    python -m binder_amd.warmstart_train label --first 1000 --count 1100 --out labels_a.npz
    (and --first 2100, --first 3200 for labels_b.npz, labels_c.npz)
    python -m binder_amd.warmstart_train train --labels labels_a.npz labels_b.npz \\
        labels_c.npz --holdout 200 --epochs 60 \\
        --out binder_amd/weights/warmstart_160x192x160.npz

label: for each seed, the 3D workload generator (12 ellipsoids, 6-voxel-RMS
DCT warp, x^0.7 intensity map, 32 bins; binder_amd.synth) makes a pair, the
plain pyramid (identity start, [8, 4, 2, 1], topology correction) registers
it on the GPU, and the label is the converged field's normalised
low-frequency DCT coefficients, 16 x 16 x 16 per component
(binder_amd.warmstart.field_coefficients).
Several worker processes share the GPU; each pair is independent.

train: regenerates every labelled pair's network input (the
half-resolution bins) on the host, keeps the last `holdout` seeds as the
evaluation set, and trains the network with Adam (lr 1e-3, batch 8) on the
mean squared error, on the GPU with the network's own kernels.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import synth
from .warmstart import NCOEF, WarmStartNet, field_coefficients, network_grid, pair_inputs

SHAPE = synth.CONFIGS["c2"]["shape"]
N_SHAPES = synth.CONFIGS["c2"]["n_shapes"]
RMS = synth.CONFIGS["c2"]["rms"]


def _say(msg):
    print(f"[warmstart] {msg}", flush=True)


def make_pair(seed):
    return synth.make_pair(SHAPE, N_SHAPES, RMS, seed)


def label_one(seed):
    """(seed, coefficients [3,C,C,C] f32 (C = NCOEF), per-level iterations,
    seconds)."""
    from .FastRegistration import FastRegistration
    pair = make_pair(seed)
    fr = FastRegistration(voxels=pair.fixed, nodes=pair.moving, **synth.REGISTRATION_KW)
    t0 = time.perf_counter()
    fr.run(list(synth.PYRAMID), synth.EM_ITERATIONS)
    dt = time.perf_counter() - t0
    coef = field_coefficients(fr.deformedVoxelPositions).astype(np.float32)
    return seed, coef, [int(lv[1]) for lv in fr.level_log], dt


def inputs_one(seed):
    """Network input bins (fixed, moving) of one pair, uint8."""
    pair = make_pair(seed)
    return seed, pair_inputs(pair.binned_fixed, pair.binned_moving, network_grid(SHAPE))


def _pool(workers):
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    # spawn: the workers start without the parent's GPU state
    return ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))


def label(first, count, out, workers=12):
    seeds = list(range(first, first + count))
    coefs = np.empty((count, 3, NCOEF, NCOEF, NCOEF), np.float32)
    iters = np.zeros((count, len(synth.PYRAMID)), np.int32)
    secs = np.zeros(count)
    t0 = time.perf_counter()
    with _pool(workers) as ex:
        for k, (seed, c, it, dt) in enumerate(ex.map(label_one, seeds, chunksize=1)):
            i = seed - first
            coefs[i], iters[i], secs[i] = c, it, dt
            if (k + 1) % 50 == 0 or k + 1 == count:
                _say(f"labelled {k + 1}/{count} pairs in {time.perf_counter() - t0:.0f}s "
                     f"(last: iterations {it}, {dt:.2f}s)")
    np.savez(out, seeds=np.array(seeds), coef=coefs, iterations=iters, seconds=secs)
    return out


def load_labels(paths):
    seeds, coefs = [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            seeds.append(z["seeds"])
            coefs.append(z["coef"])
    seeds = np.concatenate(seeds)
    coefs = np.concatenate(coefs)
    order = np.argsort(seeds, kind="stable")
    return seeds[order], coefs[order]


def regenerate_inputs(seeds, workers=16):
    grid = network_grid(SHAPE)
    fb = np.empty((len(seeds),) + grid, np.uint8)
    mb = np.empty_like(fb)
    pos = {int(s): i for i, s in enumerate(seeds)}
    t0 = time.perf_counter()
    with _pool(workers) as ex:
        for k, (seed, (f, m)) in enumerate(ex.map(inputs_one, [int(s) for s in seeds],
                                                  chunksize=4)):
            fb[pos[seed]], mb[pos[seed]] = f, m
            if (k + 1) % 200 == 0 or k + 1 == len(seeds):
                _say(f"inputs {k + 1}/{len(seeds)} in {time.perf_counter() - t0:.0f}s")
    return fb, mb


def train(label_paths, out, holdout=200, epochs=60, batch=8, lr=1e-3, seed=0,
          time_budget=None, workers=16):
    seeds, coefs = load_labels(label_paths)
    n = len(seeds)
    ntrain = n - holdout
    if ntrain < batch:
        raise ValueError("not enough labelled pairs")
    _say(f"{n} labelled pairs: {ntrain} train (seeds {seeds[0]}..{seeds[ntrain - 1]}), "
         f"{holdout} held out")
    fb, mb = regenerate_inputs(seeds, workers)
    net = WarmStartNet(network_grid(SHAPE), max_batch=batch, seed=seed, ncoef=coefs.shape[-1])
    net.set_dataset(fb, mb, coefs)
    del fb, mb
    rng = np.random.default_rng(seed)
    val = np.arange(ntrain, n)
    base = float(np.mean(coefs[val] ** 2))  # loss of predicting zero (the identity)
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(ntrain)
        steps = ntrain // batch
        idx = perm[:steps * batch].reshape(steps, batch)
        losses = net.train(idx, lr=lr)
        pv = net.predict_dataset(val)
        vloss = float(np.mean((pv - coefs[val]) ** 2))
        hist.append((ep, float(losses.mean()), vloss))
        _say(f"epoch {ep + 1}/{epochs}: train {losses.mean():.5f}, held-out {vloss:.5f} "
             f"(identity {base:.5f}), {time.perf_counter() - t0:.0f}s")
        if time_budget and time.perf_counter() - t0 > time_budget:
            _say("time budget reached")
            break
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    net.save(out, train_seeds=seeds[:ntrain], holdout_seeds=seeds[ntrain:],
             history=np.array(hist), batch=batch, lr=lr, identity_loss=base)
    _say(f"saved {out}")
    return net, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("label")
    a.add_argument("--first", type=int, required=True)
    a.add_argument("--count", type=int, required=True)
    a.add_argument("--out", required=True)
    a.add_argument("--workers", type=int, default=12)
    t = sub.add_parser("train")
    t.add_argument("--labels", nargs="+", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--holdout", type=int, default=200)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--time-budget", type=float, default=None)
    t.add_argument("--workers", type=int, default=16)
    args = ap.parse_args(argv)
    if args.cmd == "label":
        label(args.first, args.count, args.out, args.workers)
    else:
        train(args.labels, args.out, args.holdout, args.epochs, args.batch, args.lr,
              args.seed, args.time_budget, args.workers)


if __name__ == "__main__":
    sys.exit(main())
