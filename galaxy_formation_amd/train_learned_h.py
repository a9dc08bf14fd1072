"""Trains the learned first-radius network on snapshots the simulation itself
produces -- the rotating clouds (5,000 + 5,000 and 100,000 + 100,000
particles) and a smaller disk-in-halo state (100,000 + 100,000, another seed
than any benchmark input) -- on the GPU, with libgf.so's forward, backward
and Adam kernels.

    python -m galaxy_formation_amd.train_learned_h [--out PATH] [--samples N]

Writes the parameters, feature statistics and the calibrated margin (.npz)
and prints a JSON summary.
"""

import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "learned_h_k64.npz")


def build_dataset(model, runs, max_gap=4, samples=2_000_000, seed=0):
    import torch

    from . import ics
    from .stepper import Stepper, training_pairs
    recs = []
    for name, n_gas, seed_ic, steps, dt_factor in runs:
        make = ics.make_disk if name.startswith("disk") else ics.make_cloud
        p = make(n_gas, n_gas, seed=seed_ic)
        S = Stepper(p, k=model_k(model))
        S.run(steps, dt_factor=dt_factor)
        recs.append(S.records)
    # raw features first (for the statistics), then normalised ones
    k = model_k(model)
    raw = torch.cat([training_pairs(r, model, k, max_gap)[1] for r in recs])
    model.fit_norm(raw)
    parts = [training_pairs(r, model, k, max_gap) for r in recs]
    x = torch.cat([p[0] for p in parts])
    t = torch.cat([p[2] for p in parts])
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    if x.shape[0] > samples:
        keep = torch.randperm(x.shape[0], device=x.device, generator=g)[:samples]
        x, t = x[keep].contiguous(), t[keep].contiguous()
    return x, t


def model_k(model):
    return getattr(model, "k", 64)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--quantile", type=float, default=0.999)
    a = ap.parse_args(argv)
    import torch

    from .learned_h import LearnedH
    t0 = time.time()
    model = LearnedH()
    model.k = 64
    # (label, gas particles, IC seed, leapfrog steps, step factor); factor 1
    # is the long step of simulation.py:117 (10th percentile of the
    # admissible steps), 0.2 the scale of the sub-steps of a 6-rung ladder
    runs = [("cloud10k", 5000, 3, 40, 1.0), ("cloud10k", 5000, 4, 40, 3.0),
            ("cloud200k", 100000, 3, 10, 1.0), ("cloud200k", 100000, 5, 10, 3.0),
            ("disk200k", 100000, 17, 8, 1.0), ("disk200k", 100000, 19, 8, 2.5),
            ("cloud200k", 100000, 7, 12, 0.2), ("disk200k", 100000, 23, 12, 0.2)]
    x, t = build_dataset(model, runs, samples=a.samples)
    n = x.shape[0]
    perm = torch.randperm(n, device=x.device, generator=torch.Generator(device=x.device).manual_seed(1))
    nv = max(1, n // 10)
    xv, tv = x[perm[:nv]].contiguous(), t[perm[:nv]].contiguous()
    xt, tt = x[perm[nv:]].contiguous(), t[perm[nv:]].contiguous()
    t_data = time.time() - t0
    hist = model.train(xt, tt, steps=a.train_steps, log_every=max(1, a.train_steps // 10),
                       tau=a.quantile)
    t_train = time.time() - t0 - t_data
    mse0 = float((tv.double() ** 2).mean().item())            # no correction
    mse = float(((tv - model.predict(xv)).double() ** 2).mean().item())
    margin = model.calibrate_margin(xv, tv, a.quantile)
    q0 = float(torch.quantile(tv.double(), a.quantile).item())
    model.margin_fixed = q0
    # mean overshoot of the first radius over the true d_k (log), at the same
    # one-pass fraction: network vs a fixed margin on d_k(prev)
    over_net = float((model.predict(xv) + margin - tv).double().mean().item())
    over_fix = float((q0 - tv).double().mean().item())
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    print(json.dumps({"samples": n, "train": n - nv, "val": nv, "val_mse": mse,
                      "val_mse_without_network": mse0, "margin": margin,
                      "margin_without_network": q0, "mean_log_overshoot": over_net,
                      "mean_log_overshoot_without_network": over_fix, "loss_history": hist,
                      "seconds_data": t_data, "seconds_train": t_train, "out": a.out}))


if __name__ == "__main__":
    main()
