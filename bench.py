#!/usr/bin/env python
"""Accel+Potential throughput on a 1e7-particle Plummer sphere, one MI355X.

Workload (one "step"): the self-gravity Accel+Potential evaluation the reference's benchmarks time
(examples/benchmark_scaling.py:104-107 times a whole call, tree build included): bounding cube,
FP64 Morton keys, radix sort, octree build, FP64 moments (with quadrupoles when the criterion needs
them), FP32 walk pack, cell-aligned target groups, and one fused potential+acceleration grouped walk
in FP32 (node-relative coordinates, FP32 accumulation), results scattered back to input order.
The headline configuration uses the learned opening criterion: a node is accepted beyond the
opening radius the shipped MLP predicts for a per-node tolerance of alpha x the RMS field (the RMS
field is estimated per step by a 4096-target sample walk; the radii are computed per step by the
fused inference pass), with the reference's geometric test kept as the guard at theta=2, and the
group expansion: an accepted element whose third-order Taylor expansion about the target group's
centre is within expansion x the RMS field is folded into that expansion (evaluated once per
target) instead of being evaluated per target, provided that the expansion's truncation bound is also
below expansion_rel x the element's own field there (a cap on the group radius over the distance,
which keeps the worst-case error down).  The defaults (alpha 5e-4, expansion 4e-3, expansion_rel 3)
give acceleration RMS and max errors and a potential RMS error at or below the plain theta=0.7
walk's, which the same run also times ("plain_theta_0.7") on the same targets.  Inputs (positions, masses, softenings) are resident in HBM
before the timed region.

Execution (--mode): the headline ("graph") captures the whole step once with Forest.capture_step --
a sync-free build into arrays sized from an eager build of the same particles, the groups, the field
scale, the radii and the walk -- and times replays of that one HIP graph; every replay recomputes
everything from the positions in HBM and reads the build's counts back once (overflow check).  Its
output is checked bit-equal to the eager step's and its errors are scored on it.  The eager step
(--mode eager: launches from the host, host reads inside the build) is timed in the same run, and
the phase breakdown (build alone, walk by HIP events) comes from it.  The plain theta=0.7 step is
timed the same way as the headline.  Particles: Plummer sphere, M=1, a=1, truncated at 100 scale radii,
h=0.01, seed 0.

Accuracy gates, against FP64 direct summation on 1e4 random targets with the error definitions of
examples/error_benchmark.py:75-98: acceleration RMS error <= 1.8e-3 of the RMS field (the
reference's theta=0.7 figure, README.md:124), and ("gate_vs_plain_theta_0.7") acceleration RMS,
acceleration max and potential RMS errors each at most the plain theta=0.7 walk's.

Prints one JSON line (rank 0).  ``roofline`` describes the walk kernel (FP32 VALU-bound):
algorithmic flops = 20 per target-element interaction (monopole potential+acceleration; the
interaction count comes from the kernel's own counters) divided by the walk's average duration
measured with HIP events on the launch stream.  ``cpu_baseline`` times the CPU oracle (a C
restatement of the reference algorithm: serial radix build + OpenMP grouped FP64 walk at the
reference default group_size=8) on a bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOPS_PER_INTERACTION = 20  # monopole potential + acceleration
FLOPS_PER_FOLD = 60  # one element added to a group's third-order expansion (walk.hip exp_add)
FLOPS_PER_EXPANSION_EVAL = 100  # potential + acceleration of the expansion at one target (exp_eval)
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_COLUMN_PAIR = 10  # binned column: 3 differences, 5 for r^2, 2 divisions (M / r^2 / bin area)
FLOPS_PER_RAY_VISIT = 20  # ray walk: 3 differences, 5 for r^2, sqrt, 5 for the projection, opening test
# Published reference figure: tree Potential + Accel at N=1e7 Plummer, theta=0.7, 32 threads, two calls:
# 1.005e-6 + 1.124e-6 s/particle (examples/benchmark_scaling.json:51,131)
REF_PUBLISHED = 1.0 / (1.005e-6 + 1.124e-6)
# HBM traffic per walk launch is measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes, which
# cannot run inside this process: the bench line leaves roofline.traffic null and names the profile
# summary that holds the calibrated figure for the default configuration.
WALK_TRAFFIC_PROFILE = "profiles/k_walk_capped_plummer1e7_hbm_traffic.md"


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("-N", type=int, default=None, help="particles (default 1e7; 1e6 for --workload column)")
    ap.add_argument("--workload", choices=("plummer", "batched", "column", "targets"), default="plummer",
                    help="plummer: one N-particle Plummer sphere (the headline); batched: 512 independent "
                         "Plummer clusters (N log-uniform in [1e3, 1e5], scale radii 0.1-10) in one forest; "
                         "column: ColumnDensity of an N=1e6 Gaussian cloud (angle-binned, 6-ray walk beside it); "
                         "targets: AccelTarget+PotentialTarget of 1e6 uniform targets from N=1e7 Plummer sources")
    ap.add_argument("--criterion", choices=("error", "theta", "learned"), default="learned",
                    help="opening criterion of the timed walk: 'error' bounds each accepted node's dropped "
                         "quadrupole term by alpha x (RMS field); 'learned' accepts a node beyond the opening "
                         "radius the trained MLP predicts for that tolerance; 'theta' is the reference's "
                         "geometric test")
    ap.add_argument("--alpha", type=float, default=None,
                    help="per-node error tolerance relative to the RMS field (default 5e-4 learned, 8.5e-4 error)")
    ap.add_argument("--theta", type=float, default=None,
                    help="opening angle (criterion theta) or loose geometric cap (criterion error); default 0.7 / 2.0")
    ap.add_argument("--expansion", type=float, default=None,
                    help="group-expansion tolerance in units of the RMS field: accepted elements far enough "
                         "from a group's centre are folded into a third-order expansion about it (default 4e-3 "
                         "with the learned criterion, off otherwise)")
    ap.add_argument("--no-expansion", action="store_true", help="walk without the group expansion")
    ap.add_argument("--expansion-rel", type=float, default=None,
                    help="fold an element only if its truncation bound is also below this fraction of its own "
                         "field at the group centre (a cap on R/D; 0 = none; default 3 with the learned "
                         "criterion's expansion, which keeps the max error at or below the theta=0.7 walk's)")
    ap.add_argument("--bf16", action="store_true", help="learned criterion: bfloat16 MLP inference for the radii")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain theta=0.7 comparison line")
    ap.add_argument("--build", choices=("eager", "resident"), default="eager",
                    help="eager: every step builds with exact node counts (host reads inside the build); "
                         "resident: steps build without host synchronisation into arrays sized by the "
                         "previous build's node count, verified after the step (ResidentBuilder)")
    ap.add_argument("--mode", choices=("graph", "eager"), default="graph",
                    help="graph (headline): the whole step -- sync-free build, groups, field scale, radii, walk -- "
                         "captured once in one HIP graph and replayed per step; eager: the same step as a stream "
                         "of launches from the host (its phases are timed either way)")
    ap.add_argument("--group-size", type=int, default=64)
    ap.add_argument("--variant", type=int, default=0,
                    help="FP32 walk kernel variant (engine.WALK_VARIANTS: waves per SIMD, interaction-list length)")
    ap.add_argument("--softening", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--oracle-targets", type=int, default=10_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the whole-step HIP graph line")
    ap.add_argument("--cpu-sample-frac", type=float, default=0.125)
    ap.add_argument("--quadrupole", action="store_true", help="walk with quadrupole moments")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MiB",
                    help="only run a plain device copy of this many MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def cpu_threads():
    """Host threads for the CPU legs: the cores in this process's affinity mask, capped at the
    OMP_NUM_THREADS the environment sets (the GPU box gives one GPU's job a 16-core share and sets
    OMP_NUM_THREADS=16 however many cores the host shows)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else aff
    return min(aff, cap), aff


def cpu_baseline(x, m, h, theta, frac, seed, reps=3):
    """Oracle (C restatement of the reference) on the host cores, timed as one warm-up and then the
    median of `reps` runs: a full serial radix build of all N particles, then the grouped FP64 walk
    at the reference default group_size=8 (OpenMP over groups) on evenly spaced slices of the
    groups (aligned to whole groups, so each sampled group is exactly the group the full run would
    form), scaled to all N."""
    import oracle

    cores, aff = cpu_threads()
    n = len(x)
    nslices = 16
    gs = 8
    per = max(gs, int(n * frac / nslices) // gs * gs)
    starts = [(k * (n // nslices)) // gs * gs for k in range(nslices)]

    def once():
        t0 = time.perf_counter()
        tree = oracle.build_tree(x, m, h)
        t_build = time.perf_counter() - t0
        idx = tree.TreewalkIndices
        xs, hs = np.ascontiguousarray(x[idx]), np.ascontiguousarray(h[idx])
        t_walk, done = 0.0, 0
        for s in starts:
            e = min(s + per, n)
            t0 = time.perf_counter()
            oracle.fields_grouped(tree, xs[s:e], hs[s:e], potential=True, accel=True, group_size=gs, theta=theta,
                                  nthreads=cores)
            t_walk += time.perf_counter() - t0
            done += e - s
        return t_build, t_walk, done

    once()  # warm-up (page faults, OpenMP pool)
    runs = [once() for _ in range(reps)]
    done = runs[0][2]
    totals = [b + w * n / done for b, w, _ in runs]
    k = int(np.argsort(totals)[len(totals) // 2])
    t_build, t_walk, _ = runs[k]
    t_total = totals[k]
    return {
        "value": n / t_total,
        "unit": "particles/s",
        "cores": cores,
        "kind": "port",
        "sample": f"C oracle, median of {reps} runs after 1 warm-up: serial radix build of all {n} particles "
                  f"({t_build:.2f} s, one thread: the restatement's split is serial) + fused FP64 pot+acc grouped "
                  f"walk (group_size 8, theta {theta}) on {done} targets in {nslices} evenly spaced group-aligned "
                  f"slices ({t_walk:.2f} s on {cores} threads), walk scaled to N",
        "build_s": t_build,
        "walk_s_scaled_to_N": t_walk * n / done,
        "run_totals_s": totals,
        "affinity_cores": aff,
        "cpu_model": _cpu_model(),
    }


def _cpu_model():
    """The host CPU as /proc/cpuinfo names it (model name, family/model/stepping, cache) and the
    logical CPU count."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (t.strip() for t in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "cache size", "cpu MHz") and k not in info:
                    info[k] = v
            elif info:
                break
    except OSError:
        pass
    if not info:
        return "unknown"
    return (f"{info.get('model name', '?')} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}, cache {info.get('cache size', '?')}, "
            f"{info.get('cpu MHz', '?')} MHz; {os.cpu_count()} logical CPUs)")


def build_phases(build, n, quads, steps):
    """Per-phase time of the tree build from HIP events the library records at its phase boundaries
    (ptg_build_phase_events), averaged over `steps` builds, with each phase's algorithmic HBM bytes
    (per particle / per node, as below) and the rate they imply against the 8 TB/s peak.
    Algorithmic bytes: bounding cube 24 B per particle (FP64 positions read); keys 32 (24 read, 8
    written); radix sort 24 per pass (8-byte key and 4-byte index, read and written) x 8 passes of
    8-bit digits; octree build (shared levels, node count, node geometry and links, walk slots) 16 per
    particle + 32 per node; Morton-order gather 84 (4 + 40 read, 40 written); moments 40 per
    particle + 112 per node (+ 96 per node with quadrupoles); FP32 walk copy 40 written per element."""
    import ctypes

    import torch

    from pytreegrav_amd import _lib

    L = _lib.load()
    names = _lib.PTG_PHASES
    acc = np.zeros(len(names) - 1)
    nnodes = 0
    for _ in range(steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in names]
        for e in ev:
            e.record()  # creates the underlying event
        arr = (ctypes.c_void_p * len(ev))(*[e.cuda_event for e in ev])
        _lib.check(L.ptg_build_phase_events(arr, len(ev)), "ptg_build_phase_events")
        try:
            tree = build()
        finally:
            L.ptg_build_phase_events(None, 0)
        torch.cuda.synchronize()
        acc += [ev[i].elapsed_time(ev[i + 1]) for i in range(len(ev) - 1)]
        nnodes = tree.counts().nnodes if getattr(tree, "deferred", False) else tree.nnodes  # not the capacity
    ms = acc / steps
    t = dict(zip(names[1:], ms))
    quad_node = 96 if quads else 0
    phases = {
        "bbox": (t["bbox"], 24 * n),
        "keys": (t["keys"], 32 * n),
        "sort": (t["sort"], 24 * 8 * n),
        "octree": (t["count_nodes"] + t["nodes"] + t["links"], 16 * n + 32 * nnodes),
        "permute": (t["permute"], 84 * n),
        "moments": (t["moments"], 40 * n + (112 + quad_node) * nnodes),
        "fp32_walk_copy": (t["pack"], 40 * (n + nnodes)),
    }
    out = {}
    for k, (tms, b) in phases.items():
        gbs = b / (tms * 1e-3) / 1e9 if tms > 0 else None
        out[k] = {"ms": tms, "algorithmic_bytes": int(b), "GB_per_s": gbs,
                  "frac_of_hbm_peak": gbs / HBM_PEAK_GBS if gbs else None}
    out["host_gap_ms"] = t["tree_start"]  # build_sort's return (host sync) to build_tree's first launch
    out["nodes"] = int(nnodes)
    return out


def batched_flops(interactions, folds, n, expansion):
    f = FLOPS_PER_INTERACTION * interactions
    if expansion is not None:
        f += FLOPS_PER_FOLD * folds + FLOPS_PER_EXPANSION_EVAL * n
    return f


def run_batched(args):
    """512 independent Plummer clusters in one forest: one build, one grouping and one walk launch for
    all systems; the walk stage is also captured into a HIP graph and replayed."""
    import torch

    from pytreegrav_amd import Forest, engine
    from pytreegrav_amd.sampling import plummer_clusters

    dev = engine.device()
    systems = plummer_clusters(512, seed=args.seed)
    x = np.vstack([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    h = np.concatenate([s[2] for s in systems])
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])])
    N = len(x)
    pos, mass, soft = (torch.from_numpy(a).to(dev) for a in (x, m, h))
    crit = "learned" if args.criterion == "learned" else "theta"
    theta = args.theta if args.theta is not None else (0.7 if crit == "theta" else 2.0)
    alpha = (args.alpha or 1.6e-3) if crit == "learned" else None
    # the clusters' per-system RMS error grows faster with the fold tolerance than the single sphere's,
    # so the batched default folds at 1e-3 of each system's RMS field
    expansion = None if args.no_expansion else (args.expansion if args.expansion is not None
                                                else (1e-3 if crit == "learned" else None))
    rel = args.expansion_rel if args.expansion_rel is not None and expansion is not None else 0.0
    gs = args.group_size
    ctr = torch.zeros(4, dtype=torch.int64, device=dev)
    kw = dict(theta=theta, alpha=alpha, expansion=expansion, expansion_rel=rel)

    def step():
        F = Forest(pos, mass, soft, off, precision="float32", group_size=gs, criterion=crit)
        return F, F.evaluate(potential=True, accel=True, **kw)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F, r = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # walk stage alone: eager, and as a replayed HIP graph
    ctr.zero_()
    F.evaluate(potential=True, accel=True, counters=ctr, **kw)
    interactions, folds = int(ctr[0].item()), int(ctr[3].item())
    replay, out = F.capture(potential=True, accel=True, **kw)
    replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        replay()
    e1.record()
    torch.cuda.synchronize()
    graph_ms = e0.elapsed_time(e1) / args.steps
    graph_equal = bool(torch.equal(out[:, 0], r["potential"]) and torch.equal(out[:, 1:4], r["accel"]))
    # the whole step (deferred tree build, groups, field scales and radii, walk) as one graph; each
    # replay runs the graph and then reads the build's counts (one synchronisation per step)
    replay_step, out_step = F.capture_step(potential=True, accel=True, **kw)
    replay_step()
    torch.cuda.synchronize()
    step_equal = bool(torch.equal(out_step[:, 0], r["potential"]) and torch.equal(out_step[:, 1:4], r["accel"]))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        replay_step()
    torch.cuda.synchronize()
    step_graph_ms = (time.perf_counter() - t0) / args.steps * 1e3
    # accuracy: 16 systems, up to 1000 targets each, FP64 direct sum within the system
    rng = np.random.default_rng(7)
    errs = []
    acc_np = r["accel"].cpu().numpy()
    for s in rng.choice(len(systems), 16, replace=False):
        a, b = off[s], off[s + 1]
        idx = np.sort(rng.choice(b - a, min(1000, b - a), replace=False))
        xs, ms, hs = systems[s]
        d = engine.direct(xs[idx], hs[idx], xs, ms, hs, engine.PTG_ACC).cpu().numpy()
        # r == 0 pairs are skipped by the target form, so the target's own mass is excluded as in the walk
        e = np.sqrt(np.mean(np.sum((acc_np[a + idx] - d) ** 2, 1)) / np.mean(np.sum(d * d, 1)))
        errs.append(float(e))
    line = {
        "metric": "batched Accel+Potential particles/sec, 512 independent Plummer clusters in one forest",
        "value": N / dt,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_plummer_clusters", "systems": len(systems), "N_total": N,
                   "N_range": [int(np.diff(off).min()), int(np.diff(off).max())], "criterion": crit,
                   "theta": theta, "alpha": alpha, "group_expansion": expansion, "expansion_rel": rel,
                   "group_size": gs, "grouping": "cell-aligned per system",
                   "softening": "0.01 x scale radius", "scale_radius": "log-uniform 0.1-10"},
        "breakdown": {"walk_stage_graph_ms": graph_ms, "graph_replay_equals_eager": graph_equal,
                      "full_step_graph_ms": step_graph_ms, "full_step_graph_equals_eager": step_equal,
                      "full_step_graph_particles_per_s": N / (step_graph_ms * 1e-3),
                      "interactions_per_particle": interactions / N,
                      "folded_elements_per_particle": folds / N},
        "roofline": {"kernel": "k_walk (FP32 grouped walk)", "bound": "valu",
                     "achieved": batched_flops(interactions, folds, N, expansion) / (graph_ms * 1e-3) / 1e12,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": batched_flops(interactions, folds, N, expansion) / (graph_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                     "traffic": None, "convention": "walk stage (groups + walk) per replay"},
        "accuracy": {"accel_rms_per_system_max": max(errs), "accel_rms_per_system_median": float(np.median(errs)),
                     "systems_checked": 16},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_targets(args):
    """External targets: Plummer N=1e7 sources (truncated at 100 a, h=0.01) and 1e6 targets uniform in
    the cube [-2, 2]^3 (AccelTarget / PotentialTarget, frontend.py:296-418, 563-685).  A step is one
    whole call: source tree build, target Morton order, cell-aligned target groups and one fused
    potential+acceleration FP32 walk (targets in input order on the device).  The group size and walk
    kernel variant are autotuned once before the timed steps (group_size="auto")."""
    import torch

    from pytreegrav_amd import ConstructTree, engine
    from pytreegrav_amd.sampling import plummer_truncated

    dev = engine.device()
    N = args.N or 10_000_000
    NT = 1_000_000
    x, m = plummer_truncated(N, seed=args.seed, rmax=100.0)
    h = np.full(N, args.softening)
    xt = np.random.default_rng(args.seed + 1).uniform(-2.0, 2.0, size=(NT, 3))
    pos, mass, soft, tgt = (torch.from_numpy(a).to(dev) for a in (x, m, h, xt))
    tsoft = torch.zeros(NT, dtype=torch.float64, device=dev)
    mask = engine.PTG_POT | engine.PTG_ACC
    theta = args.theta if args.theta is not None else 0.7
    tree = ConstructTree(pos, mass, soft)
    engine._tuned.clear()
    order = engine.morton_order(tgt)
    tune = engine.autotune_group_size(tree.dev, tgt, tsoft, order, mask, theta=theta, precision="float32",
                                      grouping="cells")
    gs, variant = tune["choice"], tune.get("variant", 0)
    ctr = torch.zeros(4, dtype=torch.int64, device=dev)
    walk_ev = []

    def step(record):
        t = ConstructTree(pos, mass, soft)
        perm = engine.morton_order(tgt)
        ev = None
        if record:
            ctr.zero_()
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            walk_ev.append(ev)
        groups = engine.target_groups(tgt, perm, gs, mode="cells")
        if ev:
            ev[0].record()
        out = engine.walk(t.dev, tgt, tsoft, perm, mask, group_size=gs, theta=theta, precision="float32",
                          groups=groups, counters=ctr if record else None, variant=variant,
                          _events_end=ev[1] if ev else None)
        return out

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step(True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    walk_ms = float(np.mean([a.elapsed_time(b) for a, b in walk_ev]))
    interactions = int(ctr[0].item())
    sel = np.sort(np.random.default_rng(12345).choice(NT, size=min(args.oracle_targets, NT), replace=False))
    exact = engine.direct(tgt[torch.from_numpy(sel).to(dev)], tsoft[:len(sel)], pos, mass, soft, mask,
                          precision="float64").cpu().numpy()
    res = out.cpu().numpy()[sel]
    da = np.sqrt(np.sum((res[:, 1:4] - exact[:, 1:4]) ** 2, 1))
    scale = np.sqrt(np.mean(np.sum(exact[:, 1:4] ** 2, 1)))
    dp = res[:, 0] - exact[:, 0]
    achieved = FLOPS_PER_INTERACTION * interactions / (walk_ms * 1e-3) / 1e12
    line = {
        "metric": "AccelTarget+PotentialTarget targets/sec, 1e6 uniform targets in [-2,2]^3 from N=1e7 Plummer sources",
        "value": NT / dt,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plummer1e7_sources_uniform1e6_targets", "N_sources": N, "N_targets": NT,
                   "theta": theta, "criterion": "theta", "group_size": "auto", "grouping": "cell-aligned",
                   "walk_precision": "fp32 node-relative, fp32 accumulation", "build_precision": "fp64 keys and moments"},
        "autotune": {"choice_group_size": gs, "choice_variant": engine.WALK_VARIANTS[variant], "ms": tune["ms"],
                     "interactions": tune["interactions"]},
        "breakdown": {"walk_ms": walk_ms, "interactions_per_target": interactions / NT},
        "roofline": {"kernel": "k_walk (FP32 grouped walk)", "bound": "valu", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "convention": "20 flops per target-element interaction (monopole pot+acc)"},
        "accuracy": {"oracle": f"FP64 direct sum on the GPU, {len(sel)} random targets x {N} sources",
                     "accel_rms": float(np.sqrt(np.mean(da ** 2)) / scale), "accel_max": float(da.max() / scale),
                     "pot_rms": float(np.std(dp) / np.std(exact[:, 0])),
                     "pot_bias": float(np.mean(dp) / np.std(exact[:, 0]))},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_column(args):
    """ColumnDensity with rays=None (the reference's default, angle-binned grouped estimator) on a
    Gaussian cloud of N particles with radii 3 N^(-1/3): each step is one whole call (tree build,
    Morton order, binned walk over all N targets).  The exact 6-ray walk is timed beside it, and the C
    oracle on the host cores gives the CPU line."""
    import torch

    import oracle
    from pytreegrav_amd import ColumnDensity, ConstructTree, engine

    dev = engine.device()
    N = args.N or 1_000_000
    rng = np.random.default_rng(args.seed)
    x = rng.normal(size=(N, 3))
    m = np.full(N, 1.0 / N)
    h = np.full(N, 3.0 * N ** (-1 / 3))
    xt, mt, ht = (torch.from_numpy(a).to(dev) for a in (x, m, h))

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    dt = timed(lambda: ColumnDensity(xt, mt, ht), args.steps, args.warmup)
    dt_rays = timed(lambda: ColumnDensity(xt, mt, ht, rays=6), args.steps, args.warmup)
    # the walks alone on a resident tree, HIP events on the stream the engine launches on
    tree = ConstructTree(xt, mt, ht)
    tgt = tree.dev.coordinates[:N]  # the Morton-ordered positions the frontend walks
    rays = np.vstack([np.eye(3), -np.eye(3)])
    cb = torch.zeros(2, dtype=torch.int64, device=dev)
    cr = torch.zeros(1, dtype=torch.int64, device=dev)
    engine.column_binned(tree.dev, tgt, 16, 0.5)
    engine.column_rays(tree.dev, tgt, rays)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    engine.column_binned(tree.dev, tgt, 16, 0.5, counters=cb)
    e[1].record()
    e[2].record()
    engine.column_rays(tree.dev, tgt, rays, counters=cr)
    e[3].record()
    torch.cuda.synchronize()
    binned_ms, rays_ms = e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])
    pairs, visits, ray_visits = int(cb[0].item()), int(cb[1].item()), int(cr[0].item())
    achieved = pairs * FLOPS_PER_COLUMN_PAIR / (binned_ms * 1e-3) / 1e12
    rays_achieved = ray_visits * FLOPS_PER_RAY_VISIT / (rays_ms * 1e-3) / 1e12
    # CPU: the C oracle on the host cores, whole job for the binned estimator (build + all targets),
    # a 20000-target sample for the rays
    cores, _ = cpu_threads()
    t0 = time.perf_counter()
    o = oracle.build_tree(x, m, h)
    t_build = time.perf_counter() - t0
    xs = np.ascontiguousarray(x[o.TreewalkIndices])
    t0 = time.perf_counter()
    oracle.column_binned(o, xs, 16, 0.5, nthreads=cores)
    t_binned = time.perf_counter() - t0
    k = min(N, 20000)
    t0 = time.perf_counter()
    oracle.column_rays(o, x[:k], rays, nthreads=cores)
    t_rays = time.perf_counter() - t0
    line = {
        "metric": "ColumnDensity (rays=None, angle-binned) targets/sec, N=1e6 Gaussian cloud, all particles as targets",
        "value": N / dt,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "column_binned_gauss1e6", "N": N, "distribution": f"unit Gaussian, seed {args.seed}",
                   "radii": "3 N^(-1/3)", "theta": 0.5, "group_size": 16, "rays": None},
        "breakdown": {"binned_walk_ms": binned_ms, "pairs_per_target": pairs / N,
                      "visits_per_group": visits / -(-N // 16),
                      "rays6_ms_per_call": dt_rays * 1e3, "rays6_targets_per_s": N / dt_rays,
                      "rays6_walk_ms": rays_ms, "ray_visits_per_walk": ray_visits / (6 * N),
                      "rays6_walk_roofline": {"kernel": "k_column_rays", "bound": "latency (dependent link loads)",
                                              "achieved": rays_achieved, "peak": FP64_PEAK_TFLOPS,
                                              "unit": "TFLOP/s", "frac": rays_achieved / FP64_PEAK_TFLOPS,
                                              "convention": f"{FLOPS_PER_RAY_VISIT} FP64 flops per element visited"}},
        "roofline": {"kernel": "k_column_binned_wave (FP64 grouped binned walk)", "bound": "valu",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "convention": f"{FLOPS_PER_COLUMN_PAIR} FP64 flops per element-target pair (counted by the kernel)"},
        "cpu_baseline": {"value": N / (t_build + t_binned), "unit": "targets/s", "cores": cores, "kind": "port",
                         "sample": f"C oracle: serial build of all {N} particles ({t_build:.2f} s) + binned grouped walk "
                                   f"(group_size 16, theta 0.5) over all {N} targets ({t_binned:.2f} s on {cores} threads); "
                                   f"6-ray walk on {k} targets: {k / t_rays:.3g} targets/s",
                         "rays6_targets_per_s": k / t_rays, "cpu_model": _cpu_model()},
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding of the walk is a "
              f"design (targets split across GPUs with a replicated tree), not something this program launches.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        a = torch.empty(args.copy_probe * 2 ** 20, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel()}))
        return

    from pytreegrav_amd import ConstructTree, engine
    from pytreegrav_amd.sampling import plummer_truncated

    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "column":
        return run_column(args)
    if args.workload == "targets":
        return run_targets(args)
    dev = engine.device()
    N = args.N or 10_000_000
    x, m = plummer_truncated(N, seed=args.seed, rmax=100.0)
    h = np.full(N, args.softening)
    pos = torch.from_numpy(x).to(dev)
    mass = torch.from_numpy(m).to(dev)
    soft = torch.from_numpy(h).to(dev)
    mask = engine.PTG_POT | engine.PTG_ACC
    ctr = torch.zeros(4, dtype=torch.int64, device=dev)

    exact = None
    sel = None
    if not args.no_accuracy:
        rng = np.random.default_rng(12345)
        sel = np.sort(rng.choice(N, size=min(args.oracle_targets, N), replace=False))
        sel_t = torch.from_numpy(sel).to(dev)
        exact = engine.direct(pos[sel_t], soft[sel_t], pos, mass, soft, mask, self_mode=False,
                              precision="float64").cpu().numpy()

    model = None
    if args.criterion == "learned":
        from pytreegrav_amd.learned import OpeningMLP

        model = OpeningMLP.load()

    def run_config(criterion, theta, alpha, quad_walk=False, expansion=None):
        """Time `steps` whole evaluations (build + pack + groups + walk) of one configuration; the
        HIP events bracket the walk launches alone (k_walk_init + k_walk)."""
        build_quads = quad_walk or criterion in ("error", "learned")
        walk_times = []
        stats = {}
        # resident: each step's tree is built without host synchronisation into arrays sized by the
        # previous step's node count (ptg_build_deferred), and verified after the step
        builder = engine.ResidentBuilder(quadrupole=build_quads) if args.build == "resident" else None
        if builder is not None:
            builder.verify(builder(pos, mass, soft))  # the first build sizes the capacities

        def make_tree():
            return builder(pos, mass, soft) if builder is not None else \
                ConstructTree(pos, mass, soft, quadrupole=build_quads).dev

        last = {}

        def step(record):
            tree = make_tree()
            last["tree"] = tree
            ev = None
            if record:
                ctr.zero_()
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                walk_times.append(ev)
            out = engine.self_fields(tree, pos, soft, mask, theta=theta, quadrupole=quad_walk,
                                      group_size=args.group_size, criterion=criterion,
                                      alpha=alpha if criterion != "theta" else None, model=model,
                                      counters=ctr if record else None, walk_events=ev, expansion=expansion,
                                      bf16=args.bf16,
                                      stats=stats if record else None,
                                      expansion_rel=args.expansion_rel if expansion is not None else 0.0,
                                      variant=args.variant)
            if builder is not None:
                builder.verify(tree)  # never over capacity here: the positions do not change
            return out

        for _ in range(args.warmup):
            step(False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = step(True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        walk_ms = float(np.mean([a.elapsed_time(b) for a, b in walk_times]))
        # walk-only with the tree resident: the same evaluation repeated on the last step's tree, only
        # its walk launch timed (the step's own walk follows a fresh build and is a little slower)
        res_ev = []
        for _ in range(args.steps):
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            res_ev.append(ev)
            engine.self_fields(last["tree"], pos, soft, mask, theta=theta, quadrupole=quad_walk,
                               group_size=args.group_size, criterion=criterion,
                               alpha=alpha if criterion != "theta" else None, model=model, walk_events=ev,
                               expansion=expansion, bf16=args.bf16,
                               expansion_rel=args.expansion_rel if expansion is not None else 0.0,
                               variant=args.variant)
        torch.cuda.synchronize()
        walk_resident_ms = float(np.median([a.elapsed_time(b) for a, b in res_ev]))
        last.clear()
        r = {"dt": dt, "walk_ms": walk_ms, "walk_resident_ms": walk_resident_ms,
             "interactions": int(ctr[0].item()), "visits": int(ctr[1].item()),
             "particle_interactions": int(ctr[2].item()), "folds": int(ctr[3].item()),
             "groups": int(stats["groups"].item())}
        # tree build alone (same inputs), for the breakdown
        torch.cuda.synchronize()
        tb0 = time.perf_counter()
        for _ in range(args.steps):
            make_tree()
        torch.cuda.synchronize()
        r["build_ms"] = (time.perf_counter() - tb0) / args.steps * 1e3
        r["build_phases"] = build_phases(make_tree, N, build_quads, args.steps)
        r["accuracy"] = accuracy(out) if exact is not None else None
        r["out"] = out
        return r

    def run_graph_step(criterion, theta, alpha, expansion, eager_out):
        """The same evaluation as one HIP graph per step: Forest.capture_step on the single system
        (sync-free build sized by the eager tree, cell groups, field scale from the same 64 sample
        groups, opening radii, walk), replayed on the resident inputs; each replay reads the build's
        counts once afterwards (one host synchronisation per step).  Returns (summary, output)."""
        from pytreegrav_amd import Forest

        F = Forest(pos, mass, soft, [0, N], precision="float32", group_size=args.group_size,
                   criterion=criterion, field_samples=64)
        kw = dict(theta=theta, expansion=expansion,
                  expansion_rel=args.expansion_rel if expansion is not None else 0.0)
        if criterion == "learned":
            kw.update(alpha=alpha, bf16=args.bf16)
        replay, gout = F.capture_step(potential=True, accel=True, **kw)
        for _ in range(max(1, args.warmup)):
            replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            replay()
        torch.cuda.synchronize()
        gdt = (time.perf_counter() - t0) / args.steps
        return {"dt": gdt, "ms_per_step": gdt * 1e3, "particles_per_s": N / gdt,
                "equals_eager": bool(torch.equal(gout, eager_out)),
                "what": "build (sync-free, node capacity 1.25x the eager tree's), groups, field scale, radii "
                        "and walk of the same configuration captured in one HIP graph and replayed"}, gout

    def accuracy(out):
        res = out.index_select(0, torch.from_numpy(sel).to(dev)).cpu().numpy()
        a_ex, p_ex = exact[:, 1:4], exact[:, 0]
        a_t, p_t = res[:, 1:4], res[:, 0]
        da = np.sqrt(np.sum((a_t - a_ex) ** 2, axis=1))
        scale = np.sqrt(np.mean(np.sum(a_ex ** 2, axis=1)))
        dp = (p_t - p_ex)
        acc = {
            "oracle": "FP64 LDS-tiled direct sum on the GPU, target form (r==0 skipped), "
                      f"{len(sel)} random targets x {N} sources",
            "accel_rms": float(np.sqrt(np.mean(da ** 2)) / scale),
            "accel_max": float(da.max() / scale),
            # the potential's zero point is arbitrary (error_benchmark.py:90-98 scores std(dp)); its
            # constant offset is reported beside it
            "pot_rms": float(np.std(dp) / np.std(p_ex)),
            "pot_max": float(np.abs(dp - dp.mean()).max() / np.std(p_ex)),
            "pot_bias": float(np.mean(dp) / np.std(p_ex)),
            "gate_accel_rms": 1.8e-3,
        }
        acc["gate_pass"] = acc["accel_rms"] <= acc["gate_accel_rms"]
        return acc

    def gate_vs_plain(acc, ref):
        """Equal or lower error than the plain theta=0.7 walk on the same targets: accel RMS, accel max
        and potential RMS all at most the plain walk's."""
        keys = ("accel_rms", "accel_max", "pot_rms")
        out = {f"{k}_ratio_to_plain": acc[k] / ref[k] for k in keys}
        out["pass"] = all(acc[k] <= ref[k] for k in keys)
        return out

    if args.alpha is None:
        args.alpha = 5e-4 if args.criterion == "learned" else 8.5e-4
    if args.no_expansion:
        args.expansion = None
    elif args.expansion is None and args.criterion == "learned" and not args.quadrupole:
        args.expansion = 4e-3
    if args.expansion_rel is None:
        args.expansion_rel = 3.0 if args.expansion is not None and args.criterion == "learned" else 0.0
    quad = bool(args.quadrupole)
    theta = args.theta if args.theta is not None else (0.7 if args.criterion == "theta" else 2.0)
    graph_mode = args.mode == "graph"
    if graph_mode and (args.criterion not in ("learned", "theta") or quad):
        raise SystemExit("--mode graph captures the monopole theta or learned step; use --mode eager")
    main = run_config(args.criterion, theta, args.alpha, quad, args.expansion)
    graph = None
    graph_error = None
    if graph_mode or (not args.no_graph and args.criterion in ("learned", "theta") and not quad):
        try:
            graph, gout = run_graph_step(args.criterion, theta, args.alpha, args.expansion, main["out"])
        except Exception as e:  # the line still reports the eager step, and says why
            graph_error = f"{type(e).__name__}: {e}"
            print(f"bench.py: the captured step failed ({graph_error}); reporting the eager step",
                  file=sys.stderr, flush=True)
            graph_mode = False
        else:
            if graph_mode:  # the headline's errors are those of the output the timed replays wrote
                main["accuracy"] = accuracy(gout) if exact is not None else None
            del gout
    plain = None
    if not args.no_plain and not (args.criterion == "theta" and theta == 0.7 and not quad):
        p = run_config("theta", 0.7, 0.0, False)
        p_dt, p_acc, p_eq, p_err = p["dt"], p["accuracy"], None, None
        if graph_mode:
            try:
                pg, pout = run_graph_step("theta", 0.7, 0.0, None, p["out"])
            except Exception as e:  # timed eagerly instead, and the line says so
                p_err = f"{type(e).__name__}: {e}"
                print(f"bench.py: the captured theta=0.7 step failed ({p_err}); timing it eagerly",
                      file=sys.stderr, flush=True)
            else:
                p_dt, p_eq = pg["dt"], pg["equals_eager"]
                p_acc = accuracy(pout) if exact is not None else None
                del pout
        plain = {"criterion": "theta", "theta": 0.7, "moments": "monopole", "ms_per_step": p_dt * 1e3,
                 "particles_per_s": N / p_dt, "eager_ms_per_step": p["dt"] * 1e3, "graph_equals_eager": p_eq,
                 "graph_error": p_err,
                 "walk_ms": p["walk_ms"], "tree_build_ms": p["build_ms"],
                 "interactions_per_particle": p["interactions"] / N, "accuracy": p_acc,
                 "walk_speedup_of_headline": p["walk_ms"] / main["walk_ms"],
                 "step_speedup_of_headline": p_dt / (graph["dt"] if graph_mode else main["dt"])}
        if main["accuracy"] is not None and p_acc is not None:
            main["accuracy"]["gate_vs_plain_theta_0.7"] = gate_vs_plain(main["accuracy"], p_acc)
    dt = graph["dt"] if graph_mode else main["dt"]
    walk_ms, build_ms = main["walk_ms"], main["build_ms"]
    interactions, visits, acc = main["interactions"], main["visits"], main["accuracy"]
    pinter, folds = main["particle_interactions"], main["folds"]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, m, h, 0.7, args.cpu_sample_frac, args.seed)

    # counters hold the last timed step
    flops = FLOPS_PER_INTERACTION * interactions
    if args.expansion is not None:
        flops += FLOPS_PER_FOLD * folds + FLOPS_PER_EXPANSION_EVAL * N
    achieved = flops / (walk_ms * 1e-3) / 1e12
    line = {
        "metric": "Accel+Potential particles/sec at N=1e7 Plummer, equal RMS force error vs FP64 direct",
        "value": N / dt,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (N / dt) / REF_PUBLISHED,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plummer_self_accel_potential",
            "N": N,
            "distribution": "Plummer M=1 a=1 truncated at 100 a, seed %d" % args.seed,
            "softening": args.softening,
            "criterion": args.criterion,
            "theta": theta,
            "group_expansion": args.expansion,
            "expansion_rel": args.expansion_rel if args.expansion is not None else None,
            "radius_inference": ("bf16" if args.bf16 else "fp32") if args.criterion == "learned" else None,
            "alpha": args.alpha if args.criterion in ("error", "learned") else None,
            "model": (model.meta.get("data", "") + f"; margin {model.margin:.3f}") if model is not None else None,
            "group_size": args.group_size,
            "walk_variant": engine.WALK_VARIANTS[args.variant],
            "grouping": "cell-aligned",
            "walk_precision": "fp32 node-relative, fp32 accumulation",
            "build_precision": "fp64 keys and moments",
            "execution": ("one HIP graph per step (Forest.capture_step, captured once; each replay rebuilds the tree "
                          "from the resident positions and reads its counts back once)") if graph_mode else
                         "eager launches from the host, one stream plus a side stream for the groups"
                         + (f" (the captured step failed: {graph_error})" if graph_error else ""),
            "build": "sync-free (ptg_build_deferred, node capacity 1.25x)" if graph_mode else args.build,
            "moments": ("quadrupole" if quad else "monopole") + " interactions"
                       + ("; fp64 quadrupoles built for the criterion" if args.criterion != "theta" else ""),
            "parallelism": "single GPU",
        },
        "breakdown": {
            "phases_from": "the eager steps of the same configuration: build timed alone, walk by HIP events "
                           "around its launch, the rest by difference",
            "eager_step": {"ms_per_step": main["dt"] * 1e3, "particles_per_s": N / main["dt"], "build": args.build},
            "tree_build_ms": build_ms,
            "tree_build_phases": main["build_phases"],
            "pack_groups_fieldscale_radii_ms": main["dt"] * 1e3 - build_ms - walk_ms,
            "walk_ms": walk_ms,
            "walk_only_particles_per_s": N / (walk_ms * 1e-3),
            "walk_resident_ms": main["walk_resident_ms"],
            "walk_resident_particles_per_s": N / (main["walk_resident_ms"] * 1e-3),
            "interactions_per_particle": interactions / N,
            "near_field_particle_interactions_per_particle": pinter / N,
            "target_groups": main["groups"],
            "folded_elements_per_group": (folds / main["groups"]) if args.expansion is not None else None,
            "visits_per_target": visits / N,
            "full_step_graph": graph,
        },
        "roofline": {
            "kernel": "k_walk (FP32 grouped walk)",
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "traffic_profile": WALK_TRAFFIC_PROFILE,
            "flops_per_launch": flops,
            "convention": "20 flops per target-element interaction (monopole pot+acc)"
                          + ("; 60 per element folded into a group expansion, 100 per target evaluation of it"
                             if args.expansion is not None else ""),
        },
        "accuracy": acc,
        "plain_theta_0.7": plain,
        "cpu_baseline": cpu,
        "reference_published": {"value": REF_PUBLISHED, "unit": "particles/s",
                                "source": "examples/benchmark_scaling.json:51,131 (Potential+Accel, N=1e7, 32 threads)"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
