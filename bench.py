"""Force-evaluation throughput of the MI355X path on one GPU.

A step is one full force evaluation of the workload with every particle
active and the inputs resident in HBM:
  gravity  tree build (keys, radix sort, nodes, FP64 centres of mass) and the
           grouped Barnes-Hut walk for all N particles;
  hydro    gas tree, exact k-nearest-neighbour solve (radius iteration),
           h = max(d_k / 2, eps / 2), density, pressure/viscosity forces.
Two paths are timed on the same state:
  plain     FP64 pair arithmetic; the neighbour solve starts from radii
            estimated from the gas tree's cell densities;
  fast      FP32 pair arithmetic (FP64 tree, positions and centres of mass);
            the neighbour solve starts from the learned radius proposal
            (galaxy_formation_amd/data/learned_h_k64.npz, trained on the
            rotating-cloud runs by galaxy_formation_amd.train_learned_h).
The state one long step earlier (drifted back by the 10th percentile of the
admissible time steps) supplies d_k, div v and dt for the proposal.  The
neighbour solve is also timed from the other first radii (cell densities,
d_k of the previous step, d_k of the previous step with a fixed margin).
Phases are timed with HIP events (torch.cuda.Event) on the stream the
library runs on.

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
PEAK_FP32_UNPACKED = 78.6e12       # the same without packed FP32 instructions
PEAK_FP64 = 78.6e12
# HBM bytes one launch of the FP32 walk moves on the 1M disk (by precision:
# 0 monopole, 2 quadrupole): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from
# rocprofv3 --pmc (profiles/); counters cannot be read inside a run, so the
# measured figures are carried here
WALK_TRAFFIC_BYTES = {0: 1655899819, 2: 2263004592}
# the same for one evaluation's FP32-pair SPH sums on the 1M disk's gas
# (k_pack32 + k_density32 + k_force32)
SPH_TRAFFIC_BYTES = 641058784
# the accuracy gates of the metric: RMS relative acceleration error vs the
# FP64 direct sum, and density vs the plain h solve (with h equal to it)
GATE_ACC_RMS = 1e-3
# quadrupole walk: a node whose quadrupole term is bounded below this fraction
# of the group's smallest previous-step |acc| is evaluated as a monopole
QUAD_REL_TOL = 1e-4
GATE_RHO = 1e-4
PEAK_HBM = 8.0e12
GAMMA, ALPHA, BETA = 5.0 / 3.0, 1.0, 2.0
PHASES = ("tree", "walk", "gas_tree", "knn", "sph")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="disk1m")
    p.add_argument("--theta", type=float, default=0.5)
    p.add_argument("--k", type=int, default=64)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--steploop", action="store_true",
                   help="the hierarchical-leapfrog step loop instead of the force evaluation")
    p.add_argument("--substeps", type=int, default=224)
    p.add_argument("--steploop-gas-trees", default="gravity,own",
                   help="gas trees the step loop runs each path with (the faster is reported)")
    p.add_argument("--model", default=os.path.join(ROOT, "galaxy_formation_amd", "data",
                                                  "learned_h_k64.npz"))
    p.add_argument("--walk-alone", type=int, default=None, metavar="PREC",
                   help="profiling aid: build the workload's tree, then one warm-up and "
                        "three walks of precision PREC (0 FP32, 1 FP64, 2 FP32 quadrupoles), "
                        "nothing else; prints their mean duration")
    p.add_argument("--walk-blocks-beside-knn", type=int, default=0,
                   help="persistent blocks of the gravity walk in the overlapped step (0 = the "
                        "library's default, one per resident slot)")
    p.add_argument("--knn-blocks-beside-walk", type=int, default=512,
                   help="persistent blocks of the neighbour pass while the gravity walk runs "
                        "on the other stream (0 = the library default, every resident slot)")
    p.add_argument("--order", default="knn-first",
                   choices=("gas-then-walk", "knn-then-walk", "knn-first", "gas-first",
                            "walk-first"),
                   help="overlapped step: the gas tree and the neighbour solve first, the walk "
                        "after the solve (knn-then-walk); the gas passes enqueued before the "
                        "walk, beside it (knn-first); the same with the gas pipeline enqueued "
                        "before the gravity tree build (gas-first); or the walk enqueued first "
                        "(walk-first)")
    p.add_argument("--gas-tree", default="gravity", choices=("gravity", "own"),
                   help="the gas tree from the gravity tree's Morton order (gravity) or from a "
                        "sort of its own (own); orders that enqueue the gas first use their own")
    p.add_argument("--records", default="walk", choices=("walk", "build"),
                   help="pack the walk's records with the walk (walk) or right after the tree "
                        "build (build); best_schedule tries both")
    p.add_argument("--build-alone", type=int, default=None, metavar="PREC",
                   help="profiling aid: the workload's gravity tree (with the quadrupoles "
                        "for PREC 2) and gas tree, one warm-up and five builds each, "
                        "nothing else; prints their mean durations")
    p.add_argument("--knn-alone", action="store_true",
                   help="profiling aid: the gas neighbour solve alone, three times from a "
                        "first radius of 1.083 d_k (about the learned proposal's margin)")
    return p.parse_args()


def knn_alone_only(a):
    """The neighbour solve of the workload's gas and nothing else, for counter runs."""
    B = Bench(a)
    torch, L, chk = B.torch, B.L, B._lib.check
    chk(L.gf_hydro_setup_dev(B.ng, B.p(B.pos), B.p(B.vel), B.p(B.mass), B.p(B.u), B.sp))
    chk(L.gf_hydro_knn_dev(a.k, None, B.sp))
    chk(L.gf_hydro_dk_dev(B.p(B.dk_prev), B.sp))
    r0 = (B.dk_prev * 1.083).contiguous()
    chk(L.gf_hydro_knn_dev(a.k, B.p(r0), B.sp))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(B.stream)
    passes = 0
    for _ in range(3):
        passes = chk(L.gf_hydro_knn_dev(a.k, B.p(r0), B.sp))
    e1.record(B.stream)
    torch.cuda.synchronize()
    print(f"neighbour solve alone, {a.workload} gas, k = {a.k}: {e0.elapsed_time(e1) / 3:.3f} ms "
          f"per solve ({passes} pass launches)", flush=True)


def previous_state(B):
    """The state one step earlier: the same particles drifted back by a time
    step D = the 10th percentile of the particles' admissible steps (the long
    step of simulation.py:117).  Its d_k, div v and Courant steps seed the
    fast path's first radius; its FP64 gravity of every particle is the
    quadrupole walk's entry acc.  Returns D."""
    torch, c, n, ng = B.torch, B.c, B.n, B.ng
    B.step(1, None)
    torch.cuda.synchronize()
    dt_c = B.out[1]["dt"].cpu().numpy()
    ag = B.acc_g[1].cpu().numpy()
    agas = ag[:ng] + B.out[1]["acc"].cpu().numpy()
    amag = np.linalg.norm(np.concatenate([agas, ag[ng:]]), axis=1)
    dt_ind = 0.3 * np.sqrt(c.eps / np.maximum(amag, 1e-30))
    dt_ind[:ng] = np.minimum(dt_ind[:ng], dt_c)
    D = float(np.percentile(dt_ind, 10))
    pos_prev = B.pos[:ng] - D * B.vel[:ng]
    B.hydro(1, pos_prev, None)
    B._lib.check(B.L.gf_hydro_dk_dev(B.p(B.dk_prev), B.sp))
    B.divv_prev.copy_(B.out[1]["divv"])
    B.dtc_prev.copy_(B.out[1]["dt"])
    B.D = D
    acc_prev = torch.zeros_like(B.acc_g[1])
    B._lib.check(B.L.gf_tree_build_dev(n, B.p((B.pos - D * B.vel).contiguous()), B.p(B.mass),
                                       B.a.theta, B.sp))
    B._lib.check(B.L.gf_walk_dev(n, None, B.p(acc_prev), c.eps, 1, B.sp))
    torch.cuda.synchronize()
    B.acc_prev = acc_prev
    return D


def calibration_copy(B):
    """A plain copy of known size before the timed region: the reference point
    for rocprofv3's FETCH_SIZE / WRITE_SIZE counters (256 MiB read + written)."""
    torch = B.torch
    cal = torch.empty(32 << 20, dtype=torch.float64, device=B.dev)
    cal.fill_(1.0)
    cal2 = torch.empty_like(cal)
    cal2.copy_(cal)
    torch.cuda.synchronize()
    del cal, cal2


def walk_alone_only(a):
    """The gravity walk of the workload and nothing else, for counter runs."""
    B = Bench(a)
    calibration_copy(B)
    if a.walk_alone == 2:
        previous_state(B)      # the entry accelerations of the relative rule
    ms = B.walk_alone(a.walk_alone, launches=3)
    print(f"walk alone, precision {a.walk_alone}, {a.workload}: {ms:.3f} ms per launch "
          f"(mean of 3 after one warm-up)", flush=True)


def build_alone_only(a):
    """The gravity and gas tree builds of the workload and nothing else, for
    trace runs.  For PREC 2 one quadrupole walk first makes the following
    builds compute the quadrupoles with the centres of mass."""
    B = Bench(a)
    torch, L, chk = B.torch, B.L, B._lib.check
    chk(L.gf_tree_build_dev(B.n, B.p(B.pos), B.p(B.mass), a.theta, B.sp))
    if a.build_alone == 2:
        chk(L.gf_walk_dev(B.n, None, B.p(B.acc_g[2]), B.c.eps, 2, B.sp))
    ms = {"gravity tree": [], "gas tree (own sort)": [], "gas tree (from the gravity tree)": []}
    for it in range(6):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record(B.stream)
        chk(L.gf_tree_build_dev(B.n, B.p(B.pos), B.p(B.mass), a.theta, B.sp))
        e[1].record(B.stream)
        chk(L.gf_hydro_setup_dev(B.ng, B.p(B.pos), B.p(B.vel), B.p(B.mass), B.p(B.u), B.sp))
        e[2].record(B.stream)
        chk(L.gf_hydro_setup_from_tree_dev(B.ng, B.p(B.pos), B.p(B.vel), B.p(B.mass), B.p(B.u),
                                           B.sp))
        e[3].record(B.stream)
        torch.cuda.synchronize()
        if it > 0:
            ms["gravity tree"].append(e[0].elapsed_time(e[1]))
            ms["gas tree (own sort)"].append(e[1].elapsed_time(e[2]))
            ms["gas tree (from the gravity tree)"].append(e[2].elapsed_time(e[3]))
    print(f"builds alone, precision {a.build_alone}, {a.workload}: " +
          ", ".join(f"{k} {np.mean(v):.3f} ms" for k, v in ms.items()) +
          " (mean of 5 after one warm-up)", flush=True)


class Bench:
    def __init__(self, a):
        import torch

        from galaxy_formation_amd import _lib, ics
        self.torch, self._lib, self.a = torch, _lib, a
        self.L = _lib.load()
        self.dev = torch.device("cuda:0")
        c = ics.config(a.workload)
        self.c = c
        self.n = c.n
        self.gas = c.gas
        self.ng = len(self.gas)
        assert (self.gas == np.arange(self.ng)).all(), "workloads keep their gas first"
        T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(self.dev)  # noqa: E731
        self.pos, self.vel, self.mass, self.u = T(c.pos), T(c.vel), T(c.mass), T(c.u)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.dev)  # noqa: E731
        self.acc_g = {0: z(self.n, 3), 1: z(self.n, 3), 2: z(self.n, 3)}
        self.cell_rho, self.cell_count, self.r_learned = z(self.ng), z(self.ng), z(self.ng)
        self.dk_prev, self.divv_prev, self.dtc_prev, self.D = z(self.ng), z(self.ng), z(self.ng), 0.0
        self.out = {p: dict(h=z(self.ng), rho=z(self.ng), acc=z(self.ng, 3), dudt=z(self.ng),
                            dt=z(self.ng), divv=z(self.ng)) for p in (0, 1, 2)}
        # a stream of its own (default priority): the library then runs on the
        # stream the events are recorded on (handle 0 would select its own)
        self.stream = torch.cuda.Stream(device=self.dev)
        torch.cuda.set_stream(self.stream)
        self.sp = self.stream.cuda_stream
        # the gas pipeline runs on a second stream (default priority) beside
        # the gravity pipeline: the two are independent until the sum
        self.stream2 = torch.cuda.Stream(device=self.dev)
        self.sp2 = self.stream2.cuda_stream
        self.fork = torch.cuda.Event()
        self.join = torch.cuda.Event()
        self.knn_done = torch.cuda.Event()
        self.knn_go = torch.cuda.Event()
        self.tree_done = torch.cuda.Event()
        # the previous step's gravity (set by main): the quadrupole walk reads
        # it from its acc array on entry, as a simulation's acc_grav holds it
        self.acc_prev = None
        # set inside a step once its gravity tree build is enqueued (the gas
        # tree may then come from it)
        self.grav_built = False
        self.caller_radii = False

    def p(self, t):
        return None if t is None else t.data_ptr()

    def hydro(self, prec, pos, first, ev=None, rec=None, stream=None):
        """first: None (cell densities), ("radius", r) or ("learned", model)."""
        self.hydro_begin(prec, pos, first, rec, stream)
        return self.hydro_end(prec, rec, stream)

    def hydro_begin(self, prec, pos, first, rec=None, stream=None, after=None, passes=True):
        """Gas tree, first radii, and the neighbour solve's first passes
        enqueued without waiting for the device (the passes after the event
        `after`, when given; with passes=False they are left for
        hydro_passes)."""
        stream = stream or self.stream
        L, chk, sp = self.L, self._lib.check, stream.cuda_stream
        rec(2) if rec else None
        if self.grav_built and pos is self.pos and self.a.gas_tree == "gravity":
            # the gas tree from the gravity tree's Morton order (the gas is
            # the first ng rows): no sort of its own; the stream waits for
            # the gravity build's sort
            chk(L.gf_hydro_setup_from_tree_dev(self.ng, self.p(pos), self.p(self.vel),
                                               self.p(self.mass), self.p(self.u), sp))
        else:
            chk(L.gf_hydro_setup_dev(self.ng, self.p(pos), self.p(self.vel), self.p(self.mass),
                                     self.p(self.u), sp))
        rec(3) if rec else None
        r_init = None
        if first is not None and first[0] == "radius":
            r_init = first[1]
        elif first is not None and first[0] == "learned":
            model = first[1]
            chk(L.gf_hydro_cell_density_dev(self.a.k, self.p(self.cell_rho),
                                            self.p(self.cell_count), sp))
            x = model.features(self.a.k, self.dk_prev, self.cell_rho, self.cell_count,
                               self.divv_prev, self.dtc_prev, self.D, stream=stream)
            r_init = model.propose(x, self.dk_prev, out=self.r_learned, stream=stream)
        # the solve's set-up, then an event the walk can wait for, then the
        # passes: two from a caller's radii (the learned start resolves all
        # but a few groups in one), five from the cell densities (the plain
        # start's usual count)
        chk(L.gf_hydro_knn_begin_dev(self.a.k, self.p(r_init), 0, sp))
        self.caller_radii = r_init is not None
        if passes:
            self.hydro_passes(stream, after)

    def hydro_passes(self, stream, after=None):
        """The solve's first passes, after the event `after` when given; the
        event knn_go marks their start."""
        if after is not None:
            stream.wait_event(after)
        self.knn_go.record(stream)
        self._lib.check(self.L.gf_hydro_knn_more_dev(2 if self.caller_radii else 5,
                                                     stream.cuda_stream))

    def hydro_end(self, prec, rec=None, stream=None):
        """The neighbour solve's end (waits for it), then h, density, forces."""
        stream = stream or self.stream
        L, chk, sp, o = self.L, self._lib.check, stream.cuda_stream, self.out[prec]
        passes = chk(L.gf_hydro_knn_end_dev(sp))
        rec(4) if rec else None
        # the fast paths' density and forces: FP32 pair terms, FP64 sums
        chk(L.gf_set_sph_precision(0 if prec == 1 else 1))
        chk(L.gf_hydro_forces_dev(self.c.eps, GAMMA, ALPHA, BETA,
                                  *[self.p(o[x]) for x in ("h", "rho", "acc", "dudt", "dt", "divv")],
                                  sp))
        rec(5) if rec else None
        return passes

    def entry_acc(self, prec):
        """The quadrupole walk's entry accelerations: the previous step's
        gravity is put back into its acc array (on the first stream), and the
        relative rule is switched on for it alone."""
        rel = prec == 2 and self.acc_prev is not None
        self._lib.check(self.L.gf_set_quad_relative_tolerance(QUAD_REL_TOL if rel else 0.0))
        if rel:
            self.acc_g[2].copy_(self.acc_prev)

    def step(self, prec, first, ev=None, overlap=True):
        """Gravity on the first stream, the gas on the second (or, without
        overlap, after it on the first), joined at the end (events 0-1, 6 on
        the first stream, 2-5 on the gas stream)."""
        torch, L, chk, sp = self.torch, self.L, self._lib.check, self.sp
        gs = self.stream2 if overlap else self.stream
        # beside the walk the neighbour pass takes fewer persistent blocks, so
        # that the walk keeps more of each CU; alone it takes them all
        chk(L.gf_set_knn_blocks(self.a.knn_blocks_beside_walk if overlap else 0))
        # and the walk's persistent grid beside the solve (0 = the library's)
        chk(L.gf_set_walk_blocks(self.a.walk_blocks_beside_knn if overlap else 0))

        def rec(i, st):
            if ev is not None:
                ev[i].record(st)
        self.grav_built = False
        rec(0, self.stream)
        self.entry_acc(prec)          # inside the timed step
        self.fork.record(self.stream)
        if overlap and self.a.order == "gas-then-walk":
            # the whole gas pipeline first (its tree beside the gravity tree
            # build, both chains of short launches), the walk after it: the
            # two long kernels, the neighbour pass and the walk, each have
            # every CU to themselves, and no pass queues behind the walk
            gs.wait_event(self.fork)
            with torch.cuda.stream(gs):
                self.hydro_begin(prec, self.pos, first, lambda i: rec(i, gs), stream=gs)
            chk(L.gf_tree_build_dev(self.n, self.p(self.pos), self.p(self.mass), self.a.theta, sp))
            rec(1, self.stream)
            with torch.cuda.stream(gs):
                passes = self.hydro_end(prec, lambda i: rec(i, gs), stream=gs)
            self.join.record(gs)
            self.stream.wait_event(self.join)
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, sp))
            rec(6, self.stream)
            return passes
        if overlap and self.a.order == "knn-then-walk":
            # the gas tree and the neighbour solve go first, beside the
            # gravity tree build (both chains of short, latency-bound
            # launches); the walk waits for the solve, so each of the two
            # long kernels has every CU to itself; the SPH sums run beside
            # the walk
            gs.wait_event(self.fork)
            with torch.cuda.stream(gs):
                self.hydro_begin(prec, self.pos, first, lambda i: rec(i, gs), stream=gs)
            self.knn_done.record(gs)
            chk(L.gf_tree_build_dev(self.n, self.p(self.pos), self.p(self.mass), self.a.theta, sp))
            rec(1, self.stream)
            self.stream.wait_event(self.knn_done)
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, sp))
            rec(6, self.stream)
            with torch.cuda.stream(gs):
                passes = self.hydro_end(prec, lambda i: rec(i, gs), stream=gs)
            self.join.record(gs)
            self.stream.wait_event(self.join)
            return passes
        if overlap and self.a.order == "gas-first":
            # the gas pipeline's chain of short launches is enqueued first,
            # up to the neighbour solve's set-up, so it starts at once; then
            # the gravity tree build; the neighbour passes wait for the build
            # (their blocks would hold the CUs its kernels need) and the walk
            # for the passes' start
            gs.wait_event(self.fork)
            with torch.cuda.stream(gs):
                self.hydro_begin(prec, self.pos, first, lambda i: rec(i, gs), stream=gs,
                                 passes=False)
            chk(L.gf_tree_build_dev(self.n, self.p(self.pos), self.p(self.mass), self.a.theta, sp))
            if self.a.records == "build":
                chk(L.gf_walk_prepare_dev(prec, sp))
            rec(1, self.stream)
            self.tree_done.record(self.stream)
            with torch.cuda.stream(gs):
                self.hydro_passes(gs, after=self.tree_done)
            self.stream.wait_event(self.knn_go)
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, sp))
            rec(6, self.stream)
            with torch.cuda.stream(gs):
                passes = self.hydro_end(prec, lambda i: rec(i, gs), stream=gs)
            self.join.record(gs)
            self.stream.wait_event(self.join)
            return passes
        chk(L.gf_tree_build_dev(self.n, self.p(self.pos), self.p(self.mass), self.a.theta, sp))
        self.grav_built = True
        if self.a.records == "build":
            # the walk's records are packed with the build, while the gas
            # pipeline on the other stream is still short of its passes
            chk(L.gf_walk_prepare_dev(prec, sp))
        rec(1, self.stream)
        if overlap and self.a.order == "knn-first":
            # the gas tree builds beside the gravity tree; the neighbour
            # passes start once the gravity tree is built (its short kernels
            # would otherwise wait for CUs the passes hold), and the walk
            # waits for the passes' start: the pass's blocks are resident
            # when the walk's persistent grid takes the rest of each CU
            self.tree_done.record(self.stream)
            gs.wait_event(self.fork)
            with torch.cuda.stream(gs):
                self.hydro_begin(prec, self.pos, first, lambda i: rec(i, gs), stream=gs,
                                 after=self.tree_done)
            self.stream.wait_event(self.knn_go)
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, sp))
            rec(6, self.stream)
            with torch.cuda.stream(gs):
                passes = self.hydro_end(prec, lambda i: rec(i, gs), stream=gs)
        else:
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, sp))
            rec(6, self.stream)
            gs.wait_event(self.fork)
            with torch.cuda.stream(gs):
                passes = self.hydro(prec, self.pos, first, ev, lambda i: rec(i, gs), stream=gs)
        self.join.record(gs)
        self.stream.wait_event(self.join)
        return passes

    def walk_alone(self, prec, launches=5):
        """Average duration of the walk on a built tree, nothing else running
        (HIP events on the stream it is launched on, around each launch; the
        entry accelerations are put back between launches)."""
        torch, L, chk = self.torch, self.L, self._lib.check
        chk(L.gf_tree_build_dev(self.n, self.p(self.pos), self.p(self.mass), self.a.theta, self.sp))
        self.entry_acc(prec)
        chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, self.sp))
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(launches)]
        torch.cuda.synchronize()
        for e0, e1 in evs:
            self.entry_acc(prec)
            e0.record(self.stream)
            chk(L.gf_walk_dev(self.n, None, self.p(self.acc_g[prec]), self.c.eps, prec, self.sp))
            e1.record(self.stream)
        torch.cuda.synchronize()
        return float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))

    def timed(self, prec, first, steps=None, overlap=True):
        torch = self.torch
        steps = steps or self.a.steps
        for _ in range(self.a.warmup):
            self.step(prec, first, overlap=overlap)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(7)] for _ in range(steps)]
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        passes = []
        w0 = time.perf_counter()
        t0.record(self.stream)
        for k in range(steps):
            passes.append(self.step(prec, first, evs[k], overlap=overlap))
        t1.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        self._lib.check(self.L.gf_sync(self.sp))
        res = {"ms": t0.elapsed_time(t1) / steps, "wall_ms": 1e3 * wall / steps,
               "passes": int(np.max(passes))}
        # (start, end) events of each phase; gravity and gas overlap in time
        spans = {"tree": (0, 1), "walk": (1, 6), "gas_tree": (2, 3), "knn": (3, 4), "sph": (4, 5)}
        for name in PHASES:
            i, j = spans[name]
            res[name + "_ms"] = float(np.mean([e[i].elapsed_time(e[j]) for e in evs]))
        inter = np.zeros(1, np.int64)
        self._lib.check(self.L.gf_interactions_dev(self._lib.ptr(inter)))
        res["interactions"] = int(inter[0])
        tot = np.zeros(1, np.int64)
        self._lib.check(self.L.gf_hydro_passes(None, self._lib.ptr(tot)))
        res["mean_passes_per_particle"] = float(tot[0]) / self.ng
        st = np.zeros(6, np.int64)
        self._lib.check(self.L.gf_hydro_stats(st, 6))
        res["knn_groups"] = int(st[4])
        res["knn_groups_to_heap_pass"] = int(st[1])
        res["knn_groups_to_global_stack"] = int(st[2])
        return res


# persistent blocks of the neighbour pass beside the walk that each path is
# timed with: 512 leaves room on each CU for the walk's waves (the walks'
# 8 KB of LDS and at most 98 VGPRs per wave then let a pass block stay
# beside 16 of them), 0 takes every resident slot
KNN_BLOCKS_TRIED = (512, 0)
# rounds of the paired learned / fixed-margin comparison of the solve
AB_ROUNDS = 8
# where the walk's records are packed: with the walk (after the neighbour
# passes have started), or right after the tree build (beside the gas
# pipeline); the plain path's walk runs beside its neighbour passes and
# gains from the second, the fast path's waits for the passes' blocks
RECORDS_TRIED = ("walk", "build")
# the gas tree: from the gravity tree's Morton order (no sort of its own), or
# built alone; the cell-density start of the plain path reads the cells of
# the tree built alone better (its first radii overshoot less)
GAS_TREES_TRIED = ("gravity", "own")
# the walk's persistent grid beside the solve: one block per resident slot
# (0, the library's 4096), or three quarters of them, which leaves each CU
# room for a second neighbour-pass block (the quadrupole walk's 8 KB of LDS
# and 98 VGPRs a wave: 12 waves and two pass blocks fit a CU)
WALK_BLOCKS_TRIED = (0, 3072)


def best_schedule(B, prec, first):
    """The overlapped step timed with each neighbour-pass grid of
    KNN_BLOCKS_TRIED, walk grid of WALK_BLOCKS_TRIED, packing of the walk's
    records of RECORDS_TRIED and gas tree of GAS_TREES_TRIED (the same steps,
    the same work); every one tried is reported under 'schedules', and the
    fastest is timed again (its own steps, so that the reported time is not
    the best of several noisy ones) for the path's time."""
    a = B.a
    fields = ("knn_blocks_beside_walk", "walk_blocks_beside_knn", "records", "gas_tree")
    keep = tuple(getattr(a, f) for f in fields)
    tried, best = {}, None
    for gt in GAS_TREES_TRIED:
        for rec in RECORDS_TRIED:
            for wb in WALK_BLOCKS_TRIED:
                for kb in KNN_BLOCKS_TRIED:
                    cfg = (kb, wb, rec, gt)
                    for f, v in zip(fields, cfg):
                        setattr(a, f, v)
                    r = B.timed(prec, first)
                    tried[f"{a.order}, knn blocks {kb or 'all'}, walk blocks {wb or 'all'}, "
                          f"records with the {rec}, gas tree {gt}"] = r["ms"]
                    if best is None or r["ms"] < best[1]:
                        best = (cfg, r["ms"])
    for f, v in zip(fields, best[0]):
        setattr(a, f, v)
    # the chosen schedule timed again; its steps leave their outputs last
    r = B.timed(prec, first)
    r.update(dict(zip(fields, best[0])), schedules=tried, best_of_schedules_ms=best[1])
    for f, v in zip(fields, keep):
        setattr(a, f, v)
    return r


def steploop(a):
    """The step loop: a hierarchical leapfrog with 6 rungs over `substeps`
    sub-steps (a multiple of 32: whole long steps), FP32 + learned first
    radius (with quadrupoles, the path that meets the accuracy gate, and
    monopoles only) against plain FP64, with the energy and angular-momentum
    drift (stepper.RungStepper).  Prints one JSON line."""
    import torch

    from galaxy_formation_amd import ics
    from galaxy_formation_amd.learned_h import LearnedH
    from galaxy_formation_amd.stepper import RungStepper
    assert a.substeps % 32 == 0, "whole long steps of 32 sub-steps"
    c = ics.config(a.workload)
    model = LearnedH().load(a.model)
    out = {"metric": "hierarchical leapfrog sub-steps/s", "unit": "sub-steps/s",
           "workload": f"{a.workload}: {c.n} particles ({int((c.kind == 0).sum())} gas), 6 rungs",
           "substeps": a.substeps, "theta": a.theta, "k": a.k}
    runs = [(name, prec, first, gt)
            for name, prec, first in (("fast_fp32_quad", "fp32q", "learned"),
                                      ("fast_fp32", "fp32", "learned"),
                                      ("plain_fp64", "fp64", "plain"))
            for gt in a.steploop_gas_trees.split(",")]
    for name, prec, first, gt in runs:
        log = lambda m: print(f"[steploop {a.workload} {name}, gas tree {gt}] {m}",  # noqa: E731
                              file=sys.stderr, flush=True)
        log("start: forces of all particles, energy")
        S = RungStepper(c, theta=a.theta, k=a.k, precision=prec, first=first, model=model,
                        quad_relative_tolerance=QUAD_REL_TOL if prec == "fp32q" else None,
                        sph_precision=0 if prec == "fp64" else 1, gas_tree=gt)
        S.start()
        torch.cuda.synchronize()
        log(f"timed: {a.substeps} sub-steps")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record()
        for _ in range(a.substeps):
            S.substep()
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        log(f"done in {wall:.2f} s; final energy")
        E1 = S.synchronised_energy()
        E0 = S.E0
        tot0, tot1 = sum(E0[:3]), sum(E1[:3])
        af = np.array(S.active_fraction)
        r = {
            "gas_tree": gt,
            "ms_per_substep": e0.elapsed_time(e1) / a.substeps,
            "wall_ms_per_substep": 1e3 * wall / a.substeps,
            "dt_min_last_long_step": S.dt_min,
            "active_fraction": {"mean": float(af.mean()), "min": float(af.min()),
                                "max": float(af.max())},
            "gravity_evals_per_s": float(af.sum() * c.n / (e0.elapsed_time(e1) * 1e-3)),
            "knn_pass_launches_mean": float(np.mean(S.passes)),
            "energy_start": {"kinetic": E0[0], "thermal": E0[1], "potential": E0[2]},
            "energy_end": {"kinetic": E1[0], "thermal": E1[1], "potential": E1[2]},
            "energy_drift": (tot1 - tot0) / abs(tot0),
            "angular_momentum_drift": (E1[3] - E0[3]) / E0[3],
            "angular_momentum_drift_max_over_long_steps":
                float(max(abs(x - E0[3]) for x in S.L_sync) / E0[3]) if S.L_sync else None,
        }
        # each path reports its faster gas tree; every one run is listed
        out.setdefault("runs", {})[f"{name}, gas tree {gt}"] = r["ms_per_substep"]
        if name not in out or r["ms_per_substep"] < out[name]["ms_per_substep"]:
            out[name] = r
        del S
        torch.cuda.empty_cache()
    out["speedup_fast_vs_plain"] = (out["plain_fp64"]["ms_per_substep"] /
                                    out["fast_fp32_quad"]["ms_per_substep"])
    out["speedup_monopole_fast_vs_plain"] = (out["plain_fp64"]["ms_per_substep"] /
                                             out["fast_fp32"]["ms_per_substep"])
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the multi-GPU "
              "decomposition is described, not built", file=sys.stderr)
        sys.exit(2)
    if a.steploop:
        steploop(a)
        return
    if a.walk_alone is not None:
        walk_alone_only(a)
        return
    if a.knn_alone:
        knn_alone_only(a)
        return
    if a.build_alone is not None:
        build_alone_only(a)
        return
    B = Bench(a)
    torch, n, ng, c = B.torch, B.n, B.ng, B.c
    calibration_copy(B)
    D = previous_state(B)

    from galaxy_formation_amd.learned_h import LearnedH
    model = LearnedH().load(a.model)

    res = {"plain_fp64": best_schedule(B, 1, None)}
    h_plain = B.out[1]["h"].cpu().numpy()
    rho_plain = B.out[1]["rho"].cpu().numpy()
    res["fast_fp32"] = best_schedule(B, 0, ("learned", model))
    res["fast_fp32_quad"] = best_schedule(B, 2, ("learned", model))
    for name, prec in (("fast_fp32", 0), ("fast_fp32_quad", 2)):
        hf, rf = B.out[prec]["h"].cpu().numpy(), B.out[prec]["rho"].cpu().numpy()
        res[name]["h_equal_plain"] = bool(np.array_equal(hf, h_plain))
        res[name]["rho_max_rel_err_vs_plain"] = float(np.max(np.abs(rf - rho_plain) / rho_plain))
    # the same steps with gravity and gas one after the other: phase times
    # without sharing the GPU
    # (each path with the gas tree and record packing of its chosen schedule)
    serial = {}
    keep = (a.records, a.gas_tree)
    for name, prec, first in (("plain_fp64", 1, None), ("fast_fp32", 0, ("learned", model)),
                              ("fast_fp32_quad", 2, ("learned", model))):
        a.records, a.gas_tree = res[name]["records"], res[name]["gas_tree"]
        serial[name] = B.timed(prec, first, overlap=False)
    a.records, a.gas_tree = keep
    for name in serial:
        res[name]["serial"] = {k: serial[name][k] for k in ("ms",) + tuple(x + "_ms" for x in PHASES)}
    # the walk kernel alone, for its roofline
    walk_ms = B.walk_alone(0)
    walk_quad_ms = B.walk_alone(2)
    node_inter = np.zeros(1, np.int64)
    B._lib.check(B.L.gf_node_interactions_dev(B._lib.ptr(node_inter)))
    # the neighbour solve from the other first radii (FP32 path, 3 steps each,
    # serial so that the solve has the GPU to itself; the gas tree from the
    # gravity tree, as the learned start's step uses it)
    ab = {"learned": {k: serial["fast_fp32"][k] for k in ("knn_ms", "passes",
                                                          "mean_passes_per_particle")}}
    import math
    for name, first in (("cell_density", None), ("dk_prev", ("radius", B.dk_prev)),
                        ("dk_prev_fixed_margin",
                         ("radius", (B.dk_prev * math.exp(model.margin_fixed)).contiguous()))):
        r = B.timed(0, first, steps=3, overlap=False)
        ab[name] = {k: r[k] for k in ("knn_ms", "passes", "mean_passes_per_particle")}
        ab[name]["h_equal_plain"] = bool(np.array_equal(B.out[0]["h"].cpu().numpy(), h_plain))
    # the learned start against the fixed margin, alternated step by step
    # (serial steps, one timed after the warm-up each), so that a drift of
    # the box's speed falls on both alike: the mean and standard error of
    # the paired differences of the solve's time
    fixed_r = (B.dk_prev * math.exp(model.margin_fixed)).contiguous()
    diffs = []
    for _ in range(AB_ROUNDS):
        t_learned = B.timed(0, ("learned", model), steps=1, overlap=False)["knn_ms"]
        t_fixed = B.timed(0, ("radius", fixed_r), steps=1, overlap=False)["knn_ms"]
        diffs.append(t_fixed - t_learned)
    ab["paired_fixed_minus_learned_ms"] = {
        "rounds": AB_ROUNDS, "mean": float(np.mean(diffs)),
        "stderr": float(np.std(diffs, ddof=1) / math.sqrt(len(diffs))),
        "min": float(np.min(diffs)), "max": float(np.max(diffs))}

    # accuracy: gravity on a seeded 2,000-particle subset vs the FP64 direct sum
    sub = np.sort(np.random.default_rng(0).choice(n, min(n, 2000), replace=False))
    subt = torch.from_numpy(sub).to(B.dev)
    exact = torch.zeros((len(sub), 3), dtype=torch.float64, device=B.dev)
    B._lib.check(B.L.gf_direct_dev(n, B.p(B.pos), B.p(B.mass), len(sub), B.p(subt), B.p(exact),
                                   c.eps, B.sp))
    torch.cuda.synchronize()
    ex = exact.cpu().numpy()
    for name, prec in (("plain_fp64", 1), ("fast_fp32", 0), ("fast_fp32_quad", 2)):
        ap = B.acc_g[prec].cpu().numpy()[sub]
        e = np.linalg.norm(ap - ex, axis=1) / np.linalg.norm(ex, axis=1)
        res[name]["grav_rms_rel_err_vs_direct"] = float(np.sqrt(np.mean(e ** 2)))
    a32, a64 = B.acc_g[0].cpu().numpy(), B.acc_g[1].cpu().numpy()
    d = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
    res["fast_fp32"]["grav_rms_rel_diff_vs_fp64"] = float(np.sqrt(np.mean(d ** 2)))
    for name in ("fast_fp32", "fast_fp32_quad"):
        r = res[name]
        r["meets_accuracy_gates"] = bool(r["grav_rms_rel_err_vs_direct"] <= GATE_ACC_RMS and
                                         r["h_equal_plain"] and
                                         r["rho_max_rel_err_vs_plain"] <= GATE_RHO)

    accuracy = accuracy_table(B)

    # the headline: the fastest FP32 + learned-h path that meets the gates
    # (the monopole walk's error is the reference Barnes-Hut's own at this
    # theta; the quadrupole walk brings it under the gate)
    paths = {"fast_fp32": (0, walk_ms, 0), "fast_fp32_quad": (2, walk_quad_ms, int(node_inter[0]))}
    passing = [k for k in paths if res[k]["meets_accuracy_gates"]]
    head = min(passing or list(paths), key=lambda k: res[k]["ms"])
    prec_h, launch_ms, nodes_h = paths[head]
    fast, plain = res[head], res["plain_fp64"]
    inter = fast["interactions"]
    # 20 flop per entry; a quadrupole node entry costs 32 more (Q.d, d.Q.d
    # and the two extra terms)
    flops = FLOP_PER_INTERACTION * inter + 32 * nodes_h
    achieved = flops / (launch_ms * 1e-3)
    out = {
        "metric": "force-evals/s at 1M particles, FP32 + learned-h vs plain FP64 GPU, at the "
                  "stated accuracy",
        "value": n / (fast["ms"] * 1e-3),
        "unit": "force-evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": fast["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (pair arithmetic; f64 positions, tree, centres of mass, SPH sums)",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {n} particles ({ng} gas), all active, "
                               "gravity + SPH (kNN, h, density, forces)",
                   "n": n, "n_gas": ng, "theta": a.theta, "eps": c.eps, "k": a.k,
                   "path": head, "first_radius_fast": "learned proposal",
                   "quad_rule": {"relative_tolerance": QUAD_REL_TOL,
                                 "entry_acc": "previous step's FP64 gravity (every particle "
                                              "drifted back by previous_step_D), restored "
                                              "inside each timed step"},
                   "gates": {"acc_rms_vs_direct": GATE_ACC_RMS, "rho_vs_plain": GATE_RHO},
                   "meets_gates": fast["meets_accuracy_gates"]},
        # compute-bound on the FP32 vector units (VALU), published peak 157.3
        # TFLOP/s with packed FP32 (78.6 without packing, which this build
        # does not use); no MFMA: pair terms are not a contraction
        "roofline": {"bound": "fp32-vector", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": WALK_TRAFFIC_BYTES[prec_h] if a.workload == "disk1m" else None,
                     "kernel": {0: "k_walk<false, false, false> (FP32 group walk, monopoles)",
                                2: "k_walk<false, false, true> (FP32 group walk, quadrupoles)"}[prec_h],
                     "launch_ms": launch_ms, "interactions": inter, "quadrupole_entries": nodes_h,
                     "flop_per_launch": flops,
                     "unpacked_fp32_peak": PEAK_FP32_UNPACKED / 1e12,
                     "frac_of_unpacked": achieved / PEAK_FP32_UNPACKED},
        "roofline_sph": sph_roofline(ng, a.k, fast["serial"]["sph_ms"], a.workload),
        # the quadrupole walk: node entries cost 52 flop (20 + Q.d, d.Q.d and
        # the two extra terms), particle entries 20
        "roofline_monopole_walk": {"kernel": "k_walk<false, false, false>", "launch_ms": walk_ms,
                                   "achieved": 20 * res["fast_fp32"]["interactions"] / (walk_ms * 1e-3) / 1e12,
                                   "unit": "TFLOP/s", "peak": PEAK_FP32 / 1e12,
                                   "traffic": WALK_TRAFFIC_BYTES[0] if a.workload == "disk1m" else None},
        "plain_fp64": {"value": n / (plain["ms"] * 1e-3), **plain},
        "fast_fp32": res["fast_fp32"],
        "fast_fp32_quad": res["fast_fp32_quad"],
        "speedup_fast_vs_plain": plain["ms"] / fast["ms"],
        "speedup_monopole_fast_vs_plain": plain["ms"] / res["fast_fp32"]["ms"],
        "first_radius_ab": ab,
        "learned_margin": model.margin, "fixed_margin": model.margin_fixed,
        "previous_step_D": D,
        "accuracy": accuracy,
    }
    out["gas_physics"] = gas_physics(B)
    if a.workload == "disk1m":
        out["host_abi_end_to_end"] = host_abi_end_to_end(B)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(B)
    print(json.dumps(out))


def gas_physics(B, launches=20, n_new=1000, k_fb=16):
    """The sub-grid steps beside the force path, on the workload's gas, each
    timed alone with HIP events on its stream:
      * thermodynamics of every gas row (simulation.py:97-104): one pass,
        48 B per row (g, u, du/dt, rho, no_cool_until read; u written);
      * supernova feedback of n_new new stars (star_formation.py:50-65): the
        k_fb-nearest query against all gas, its tree build included (the
        reference builds a cKDTree per call), and the heating; scipy's
        cKDTree build + query on the host beside it, and whether the two
        neighbour lists agree."""
    import torch
    from scipy.spatial import cKDTree

    from galaxy_formation_amd import ics
    from galaxy_formation_amd.star_formation import E_SN_PER_MSUN
    from galaxy_formation_amd.units import MSUN_G, MYR
    L, chk, sp, ng, n = B.L, B._lib.check, B.sp, B.ng, B.n
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g = torch.arange(ng, dtype=torch.int64, device=B.dev)
    u = B.u.clone()
    ncu = torch.full((n,), -np.inf, dtype=torch.float64, device=B.dev)
    dudt, rho = B.out[1]["dudt"], B.out[1]["rho"]

    def thermo():
        chk(L.gf_thermo_dev(ng, g.data_ptr(), n, u.data_ptr(), dudt.data_ptr(), rho.data_ptr(),
                            ncu.data_ptr(), 0.0, 1e-4, ics.U_PER_K, 1e4 * ics.U_PER_K, sp))
    thermo()
    torch.cuda.synchronize()
    e0.record(B.stream)
    for _ in range(launches):
        thermo()
    e1.record(B.stream)
    torch.cuda.synchronize()
    t_thermo = e0.elapsed_time(e1) / launches
    chk(L.gf_gas_errors_dev(sp))
    gbs = 48.0 * ng / (t_thermo * 1e-3) / 1e9

    rng = np.random.default_rng(0)
    new = np.sort(rng.choice(ng, n_new, replace=False))
    new_t = torch.from_numpy(new).to(B.dev)
    q = B.pos[new_t].contiguous()
    m_new = B.mass[new_t].contiguous()
    nb = torch.empty(n_new * k_fb, dtype=torch.int64, device=B.dev)
    dist = torch.empty(n_new * k_fb, dtype=torch.float64, device=B.dev)

    def query():
        chk(L.gf_knn_query_dev(ng, B.pos.data_ptr(), n_new, q.data_ptr(), k_fb, nb.data_ptr(),
                               dist.data_ptr(), sp))
    query()
    reps = 5
    torch.cuda.synchronize()
    e0.record(B.stream)
    for _ in range(reps):
        query()
    e1.record(B.stream)
    torch.cuda.synchronize()
    t_query = e0.elapsed_time(e1) / reps
    e_per_g = 0.4 * E_SN_PER_MSUN / MSUN_G

    def heat():
        chk(L.gf_feedback_heat_dev(n_new, k_fb, nb.data_ptr(), m_new.data_ptr(), e_per_g,
                                   30.0 / MYR, n, B.mass.data_ptr(), u.data_ptr(), ncu.data_ptr(),
                                   sp))
    heat()
    torch.cuda.synchronize()
    e0.record(B.stream)
    for _ in range(reps):
        heat()
    e1.record(B.stream)
    torch.cuda.synchronize()
    t_heat = e0.elapsed_time(e1) / reps

    pg = B.pos[:ng].cpu().numpy()
    qn = q.cpu().numpy()
    w0 = time.perf_counter()
    _, nb_ref = cKDTree(pg).query(qn, k=k_fb, workers=-1)
    t_cpu = 1e3 * (time.perf_counter() - w0)
    same = bool(np.array_equal(nb.cpu().numpy().reshape(n_new, k_fb), nb_ref))
    return {
        "thermodynamics": {"rows": ng, "ms": t_thermo, "bytes_per_row": 48,
                           "achieved_GBps": gbs, "frac_of_8TBps": gbs / 8000.0},
        "feedback": {"new_stars": n_new, "k": k_fb, "gas": ng,
                     "knn_query_ms_incl_tree_build": t_query, "heat_ms": t_heat,
                     "cpu_ckdtree_build_and_query_ms": t_cpu, "cpu_threads": os.cpu_count(),
                     "neighbours_equal_ckdtree": same},
    }


def accuracy_table(B):
    """RMS relative acceleration error on the seeded 2,000-particle subsets of
    the reference fixtures (tests/golden: written by the reference
    cpp/nbody.cpp built from source, FP32 Barnes-Hut at the same theta), for
    the FP32 and FP64 walks and for the reference itself, against the FP64
    direct sum; and the RMS difference of each walk from the reference."""
    torch, L, chk = B.torch, B.L, B._lib.check
    from galaxy_formation_amd import ics
    rows = []
    for name, wl, theta in (("bh_cloud200k_t04", "cloud200k", 0.4),
                            ("bh_cloud200k_t05", "cloud200k", 0.5),
                            ("bh_cloud200k_t07", "cloud200k", 0.7),
                            ("bh_disk1m_t05", "disk1m", 0.5)):
        path = os.path.join(ROOT, "tests", "golden", name + ".npz")
        if not os.path.exists(path):
            continue
        with np.load(path, allow_pickle=False) as f:
            sub, ref = f["sub"], f["acc_sub"]
        c = ics.config(wl)
        pos = torch.from_numpy(np.ascontiguousarray(c.pos)).to(B.dev)
        mass = torch.from_numpy(np.ascontiguousarray(c.mass)).to(B.dev)
        subt = torch.from_numpy(sub.astype(np.int64)).to(B.dev)
        exact = torch.zeros((len(sub), 3), dtype=torch.float64, device=B.dev)
        chk(L.gf_direct_dev(c.n, B.p(pos), B.p(mass), len(sub), B.p(subt), B.p(exact), c.eps, B.sp))
        acc = torch.zeros((c.n, 3), dtype=torch.float64, device=B.dev)
        out = {"workload": wl, "n": c.n, "theta": theta}
        torch.cuda.synchronize()
        ex = exact.cpu().numpy()

        def rms(a, b):
            e = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
            return float(np.sqrt(np.mean(e ** 2)))
        out["reference_vs_direct"] = rms(ref, ex)
        for label, prec in (("fp32", 0), ("fp64", 1), ("fp32_quad", 2)):
            chk(L.gf_gravity_dev(c.n, B.p(pos), B.p(mass), B.p(acc), theta, c.eps, None, 1, prec,
                                 B.sp))
            torch.cuda.synchronize()
            a = acc.cpu().numpy()[sub]
            out[label + "_vs_direct"] = rms(a, ex)
            out[label + "_vs_reference"] = rms(a, ref)
        rows.append(out)
    return rows


def sph_roofline(ng, k, ms, workload):
    """The density and force passes (FP32 pair terms, the fast paths) against
    HBM: algorithmic bytes per gas particle 4k + 72 (density) and 4k + 120
    (force): the int32 neighbour column entries, the particle's own state and
    outputs; neighbour records counted once per particle, as if cached."""
    b = ng * ((4 * k + 72) + (4 * k + 120))
    achieved = b / (ms * 1e-3)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM,
            "traffic": SPH_TRAFFIC_BYTES if workload == "disk1m" and k == 64 else None,
            "bytes_algorithmic": b, "kernels": "k_pack32 + k_density32 + k_force32", "ms": ms}


def host_abi_end_to_end(B, runs=3):
    """The drop-in through the reference's host-array interfaces, as
    simulation.py:61-95 calls them: Engine.gravity (nbody_accel_active,
    acc_grav holding the previous step's gravity, as simulation.py:64 passes
    it), NeighbourFinder.find (a full search, fresh = True), h = max(d_k / 2,
    eps / 2) in numpy, and sph.sph_forces; every call copies its numpy inputs
    to the GPU and its outputs back (PCIe included).  Wall time, median of
    `runs` after one warm-up, for the plain path (FP64 gravity) and the fast
    path's gravity (FP32 + quadrupoles, relative rule); the host interface
    has no learned first radius (sph.py's signatures carry no history)."""
    import torch

    from galaxy_formation_amd import sph
    from galaxy_formation_amd.engine import Engine
    c, ng, k = B.c, B.ng, B.a.k
    torch.cuda.synchronize()
    acc_prev = B.acc_prev.cpu().numpy() if B.acc_prev is not None else np.zeros((c.n, 3))
    g = np.arange(ng, dtype=np.int64)
    gp = np.ascontiguousarray(c.pos[:ng])
    gv = np.ascontiguousarray(c.vel[:ng])
    gm = np.ascontiguousarray(c.mass[:ng])
    gu = np.ascontiguousarray(c.u[:ng])
    out = {}
    for name, prec, rel in (("plain_fp64", "fp64", 0.0), ("fast_fp32_quad", "fp32q", QUAD_REL_TOL)):
        eng = Engine(prec, quad_relative_tolerance=rel)
        nf = sph.NeighbourFinder(c.n, k=k, n_candidates=k)
        rho, dudt, dt_i = np.empty(ng), np.empty(ng), np.empty(ng)
        a_sph = np.empty((ng, 3))

        def one():
            acc = acc_prev.copy()
            t = [time.perf_counter()]
            eng.gravity(c.pos, c.mass, acc, B.a.theta, c.eps)
            t.append(time.perf_counter())
            d, nb = nf.find(g, gp, True)
            h = np.maximum(0.5 * d[:, -1], 0.5 * c.eps)
            t.append(time.perf_counter())
            sph.sph_forces(gp, gv, gm, nb, h, gu, GAMMA, ALPHA, BETA, rho, a_sph, dudt, dt_i)
            t.append(time.perf_counter())
            return np.diff(t)
        one()
        r = np.array([one() for _ in range(runs)])
        tm = r[int(np.argsort(r.sum(1))[len(r) // 2])]
        out[name] = {"ms": 1e3 * float(tm.sum()), "gravity_ms": 1e3 * float(tm[0]),
                     "neighbours_ms": 1e3 * float(tm[1]), "sph_forces_ms": 1e3 * float(tm[2]),
                     "force_evals_per_s": c.n / float(tm.sum())}
        eng.quad_relative_tolerance = 0.0
    out["note"] = (f"host arrays through the C ABI, k = {k}: nbody_accel_active, "
                   "gf_nbcache_search, gf_sph_forces; H2D and D2H copies inside the times")
    return out


def cpu_baseline(B):
    """The reference path on the host cores, one full evaluation of the same
    workload: cpp/nbody.cpp built from source (oracle/_ref; nbody_accel_active,
    all active, rebuild = 1), scipy cKDTree kNN (the reference's own library,
    sph.py:166, workers = -1), and sph_forces restated in C (the oracle; the
    reference's Numba is not installed)."""
    from scipy.spatial import cKDTree

    from oracle import oracle as O
    L = O.ref_lib()
    if L is None:
        return None
    c, ng, k = B.c, B.ng, B.a.k
    pos, mass = c.pos, c.mass
    acc = np.zeros_like(pos)
    act = np.ones(c.n, np.uint8)
    gp = np.ascontiguousarray(pos[:ng])
    gv = np.ascontiguousarray(c.vel[:ng])
    gm = np.ascontiguousarray(mass[:ng])
    gu = np.ascontiguousarray(c.u[:ng])

    def one():
        t = [time.perf_counter()]
        L.nbody_accel_active(c.n, pos, mass, acc, B.a.theta, c.eps, act, 1)
        t.append(time.perf_counter())
        d, nb = cKDTree(gp).query(gp, k=k, workers=-1)
        t.append(time.perf_counter())
        h = np.maximum(0.5 * d[:, -1], 0.5 * c.eps)
        O.sph_forces(gp, gv, gm, nb, h, gu, GAMMA, ALPHA, BETA)
        t.append(time.perf_counter())
        return np.diff(t)

    one()
    runs = np.array([one() for _ in range(5)])
    totals = runs.sum(axis=1)
    tm = runs[int(np.argsort(totals)[len(totals) // 2])]
    total = float(tm.sum())
    return {"value": c.n / total, "unit": "force-evals/s", "cores": int(L.nbody_num_threads()),
            "kind": "reference", "cpu": cpu_model(),
            "sample": f"full {B.a.workload} evaluation ({c.n} particles, {ng} gas, k={k}), "
                      "median of 5 after 1 warm-up; gravity = reference nbody.cpp, kNN = "
                      "scipy cKDTree, SPH sums = C restatement (Numba absent)",
            "build_flags": "-O3 -march=x86-64-v3 -ffast-math -fopenmp (the reference's "
                           "cpp/Makefile:5 flags, x86-64-v3 in place of -march=native: the "
                           "library is built once, before the run, for any x86-64-v3 host)",
            "phases_s": {"gravity": float(tm[0]), "knn": float(tm[1]), "sph": float(tm[2])},
            "totals_s": [float(x) for x in totals]}


def cpu_model():
    """The host CPU's model name and logical CPU count, from /proc/cpuinfo."""
    try:
        names = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                 if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else None
    except OSError:
        return None


if __name__ == "__main__":
    main()
