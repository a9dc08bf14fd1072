"""Force-evaluation throughput of the MI355X path on one GPU.

A step is one full force evaluation of the workload with every particle
active, inputs resident in HBM: tree build (keys, sort, nodes, FP64 centres of
mass) and the grouped Barnes-Hut walk for all N particles.  Timed with HIP
events (torch.cuda.Event on the stream the library runs on).

    python bench.py [--steps K] [--warmup W] [--workload disk1m] [--theta 0.5]

Prints ONE JSON line (the last line of stdout).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = 20          # Nyland, Harris & Prins (GPU Gems 3, ch. 31)
PEAK_FP32 = 157.3e12               # MI355X vector FP32, FLOP/s
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="disk1m")
    p.add_argument("--theta", type=float, default=0.5)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=1000000,
                   help="particles of the workload the CPU baseline evaluates")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the multi-GPU "
              "decomposition is described, not built", file=sys.stderr)
        sys.exit(2)
    import torch

    from galaxy_formation_amd import _lib, ics

    L = _lib.load()
    dev = torch.device("cuda:0")
    c = ics.config(a.workload)
    n = c.n
    pos = torch.from_numpy(c.pos).to(dev)
    mass = torch.from_numpy(c.mass).to(dev)
    acc = {p: torch.zeros((n, 3), dtype=torch.float64, device=dev) for p in (0, 1)}
    # a stream of its own (default priority): the library then runs on the
    # stream the events are recorded on (handle 0 would select its own stream)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    P = _lib.ptr

    def tp(t):
        return t.data_ptr()

    def step(prec, ev=None):
        if ev:
            ev[0].record(stream)
        _lib.check(L.gf_tree_build_dev(n, tp(pos), tp(mass), a.theta, sp))
        if ev:
            ev[1].record(stream)
        _lib.check(L.gf_walk_dev(n, None, tp(acc[prec]), c.eps, prec, sp))
        if ev:
            ev[2].record(stream)

    res = {}
    for prec, name in ((1, "fp64"), (0, "fp32")):
        for _ in range(a.warmup):
            step(prec)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        t0.record(stream)
        for k in range(a.steps):
            step(prec, evs[k])
        t1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        _lib.check(L.gf_sync(sp))
        ms = t0.elapsed_time(t1) / a.steps
        build_ms = np.mean([e[0].elapsed_time(e[1]) for e in evs])
        walk_ms = np.mean([e[1].elapsed_time(e[2]) for e in evs])
        inter = np.zeros(1, np.int64)
        _lib.check(L.gf_interactions_dev(P(inter)))
        res[name] = dict(ms=ms, wall_ms=1e3 * wall / a.steps, build_ms=build_ms, walk_ms=walk_ms,
                         interactions=int(inter[0]))

    # accuracy on a seeded 2,000-particle subset against the FP64 direct sum
    sub = np.sort(np.random.default_rng(0).choice(n, min(n, 2000), replace=False))
    subt = torch.from_numpy(sub).to(dev)
    exact = torch.zeros((len(sub), 3), dtype=torch.float64, device=dev)
    _lib.check(L.gf_direct_dev(n, tp(pos), tp(mass), len(sub), tp(subt), tp(exact), c.eps, sp))
    torch.cuda.synchronize()
    ex = exact.cpu().numpy()
    for name, p in (("fp64", 1), ("fp32", 0)):
        ap = acc[p].cpu().numpy()[sub]
        e = np.linalg.norm(ap - ex, axis=1) / np.linalg.norm(ex, axis=1)
        res[name]["rms_rel_err_vs_direct"] = float(np.sqrt(np.mean(e ** 2)))
    a32 = acc[0].cpu().numpy()
    a64 = acc[1].cpu().numpy()
    d = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
    res["fp32"]["rms_rel_diff_vs_fp64"] = float(np.sqrt(np.mean(d ** 2)))

    main_p = res["fp32"]
    inter = main_p["interactions"]
    achieved = FLOP_PER_INTERACTION * inter / (main_p["walk_ms"] * 1e-3)
    out = {
        "metric": "force-evals/s at 1M particles, FP32 interactions vs plain FP64 GPU",
        "value": n / (main_p["ms"] * 1e-3),
        "unit": "force-evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_p["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 positions/tree",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: gravity (tree build + walk), all {n} active",
                   "n": n, "theta": a.theta, "eps": c.eps},
        "roofline": {"bound": "valu-fp32", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "k_walk<false,false>", "interactions": inter,
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "fp64_plain": {"value": n / (res["fp64"]["ms"] * 1e-3), **res["fp64"]},
        "fp32": res["fp32"],
        "speedup_fp32_vs_fp64": res["fp64"]["ms"] / res["fp32"]["ms"],
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(c, a)
    print(json.dumps(out))


def cpu_baseline(c, a):
    """The reference engine (cpp/nbody.cpp built from source, oracle/_ref) on
    the host cores: nbody_accel_active, every particle active, rebuild = 1."""
    from oracle import oracle as O

    L = O.ref_lib()
    if L is None:
        return None
    ns = min(a.cpu_sample, c.n)
    pos = np.ascontiguousarray(c.pos[:ns])
    mass = np.ascontiguousarray(c.mass[:ns])
    acc = np.zeros_like(pos)
    L.nbody_accel_active(ns, pos, mass, acc, a.theta, c.eps, np.ones(ns, np.uint8), 1)  # warm-up
    times = []
    for _ in range(3):
        t = time.perf_counter()
        L.nbody_accel_active(ns, pos, mass, acc, a.theta, c.eps, np.ones(ns, np.uint8), 1)
        times.append(time.perf_counter() - t)
    t = float(np.median(times))
    return {"value": ns / t, "unit": "force-evals/s", "cores": int(L.nbody_num_threads()),
            "kind": "reference",
            "sample": f"first {ns} particles of {a.workload}, gravity only, median of 3 after 1 warm-up"}


if __name__ == "__main__":
    main()
