"""Benchmark of the metadynamics CV+bias+force path on one MI355X.

Default (the headline metric, config 2): atom-steps/s of Q6+Q4 CV+bias forces at 256,000
atoms.  40x40x40 FCC cells of Ar (a = 5.26 A), lower half jittered by sigma = 0.6 A
("melted"), upper half by 0.1 A; CVs Q6.AVE, Q4.AVE and the Lechner-Dellago local Q6
(N = 18 nearest within 3.9 A, list cutoff 5.9 A); DIM 1 on "v1+v2" over [0,2] with 2000
bins, well-tempered (biasf 10), sigma 0.01, height 0.005, PACE 10.  A step = selection +
q_lm + local Q + FP64 reductions + hill deposit/bias evaluation + chain rule + exact
forces, one HIP graph launch per step.  Inputs are resident in HBM before the timed
region; MD integration and neighbour-list builds are outside the path and not timed.

Other workloads (each prints its own JSON line): --config c1 (4,000-atom CPU reference
case), c3 (500,000-atom structure factor), c4 (64 walkers), c5 (3-D learned bias vs exact
hills sum).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "metad-toy_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, ROOT)

from mdt.workloads import (A, C1_ARGS, C2_ARGS, C3_ARGS, C4_ARGS, C5_BASE, KB, LQ_ARGS, RC, RG,  # noqa: E402
                           c1_system, c2_system, c3_system, c4_systems, wt_hills)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# algorithmic work (FMA = 2 FLOP, transcendental = 1)
FLOP_FWD_PAIR = 261.0        # q_lm of {Q4, Q6} per selected pair
FLOP_GRAD_PAIR = 903.0       # gradient of {Q4, Q6} per selected pair
FLOP_LD_FWD_PAIR = 28.0      # local Q6 average per selected pair
FLOP_SK = 24.0               # structure factor per (atom, k): 10 forward + 14 gradient
FP64_PEAK_TFLOPS = 78.6
# STEINH L 6 per counted list entry, FP64: sigma, sigma', Y_6m and its accumulation (k_lq_q,
# 140); the weighted q average (k_lq_lq, 28); the weighted K sum and <K_j, q_a> (k_lq_h, 56);
# the pair gradient <H, Y>, grad <H, Y> and its assembly (k_lq_pair, 282)
FLOP_LQ_Q, FLOP_LQ_LQ, FLOP_LQ_H, FLOP_LQ_PAIR = 140.0, 28.0, 56.0, 282.0
LQ_PAIR_ENTRY_B, LQ_PAIR_ATOM_B = 60.0, 180.0   # k_lq_pair algorithmic bytes per list entry, per atom
METRIC = "atom-steps/s, Q6+Q4 CV+bias forces @256k atoms; NN bias vs hills @<=0.1 kT"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_run(fix, step0, steps, per_launch=None):
    """Wall time of `steps` steps on the device (graphs of `per_launch` captured steps
    per launch, the library's batch by default), ended by a state read."""
    fix.sync()
    t0 = time.perf_counter()
    fix.run(step0, steps, per_launch)
    fix.sync()
    return time.perf_counter() - t0


def timed_modes(fix, step0, steps, modes):
    """Times `steps` steps in each run mode of `modes` (captured steps per graph launch, 0
    for plain kernel launches), one after another from step0; returns (best seconds, its
    mode, {mode: seconds})."""
    ts = {}
    for k, m in enumerate(modes):
        ts[m] = timed_run(fix, step0 + k * steps, steps, m)
    best = min(ts, key=ts.get)
    return ts[best], best, ts


def mode_name(m):
    return "plain kernel launches" if m == 0 else f"{m} captured step(s) per graph launch"


def timed_plain(fix, step0, steps):
    """Wall time of `steps` host-driven steps (plain launches, one post_force call and
    state read per step)."""
    fix.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        fix.post_force(step0 + k)
    fix.sync()
    return time.perf_counter() - t0


def oracle_steps(fn, budget_s):
    """Run fn() until budget_s of CPU time is used (at least once); returns (count, seconds)."""
    k, t0 = 0, time.perf_counter()
    while True:
        fn()
        k += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            return k, dt


def steinhardt_roofline(fix, system, kms, fused_ld):
    """Roofline of the dominant per-atom kernel from the algorithmic work per launch.
    fused_ld: an unbiased local-Q CV is accumulated by the force pass (no k_ld launch)."""
    n = system.n_local
    nsel = fix.extract("Q6.nsel", n)
    pairs = float(nsel.sum())
    x, off, idx, own = system.arrays()
    c = np.diff(off).astype(np.float64)
    n_ghost = x.shape[0] - n
    bytes_sel = float(np.sum(24 + 4 * (c + 1) + 4 * nsel + 4) + 24 * n_ghost)
    work = {"k_env": FLOP_FWD_PAIR * pairs, "k_force": FLOP_GRAD_PAIR * pairs, "k_ld": FLOP_LD_FWD_PAIR * pairs}
    fused_ld = fused_ld and "k_ld" not in kms
    if fused_ld:
        # the unbiased local Q6 is accumulated by the force pass: its work moves there
        work["k_force"] += FLOP_LD_FWD_PAIR * pairs
    tms = {k: kms[k] for k in work if k in kms}
    dom = max(tms, key=lambda k: tms[k])
    t = tms[dom] * 1e-3
    rate = work[dom] / t / 1e12
    per = {"k_env": FLOP_FWD_PAIR, "k_force": FLOP_GRAD_PAIR + (FLOP_LD_FWD_PAIR if fused_ld else 0.0),
           "k_ld": FLOP_LD_FWD_PAIR}[dom]
    name = dom
    roof = {"kernel": name, "bound": "fp32", "achieved": rate, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": rate / FP32_PEAK_TFLOPS, "traffic": None,
            # code built without packed FP32 instructions issues one wave64 FP32 op per clock per
            # CU: half the listed peak
            "peak_without_packed_fp32": FP32_PEAK_TFLOPS / 2, "frac_without_packed_fp32": rate / (FP32_PEAK_TFLOPS / 2),
            "algorithmic": f"{per:.0f} FLOP x {pairs:.0f} selected pairs per step", "mean_ms": tms[dom],
            "selection_hbm_GBs": bytes_sel / (kms["k_env"] * 1e-3) / 1e9 if "k_env" in kms else None,
            "selection_bytes": bytes_sel}
    tp = os.path.join(ROOT, "profiles", f"traffic_{dom}_c2.json")
    if os.path.exists(tp):
        roof["traffic"] = json.load(open(tp)).get("bytes_per_launch")
    return roof, pairs


def line(value, steps, warmup, ms, dtype, config, roof, cpu, metric=METRIC, unit="atom-steps/s"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded lattices / uniform liquids, see config)",
            "config": config, "roofline": roof, "cpu_baseline": cpu}


class EngineBuffer:
    """Device copy of an array made with the HIP runtime the library links: the MD
    engine's own position buffer in a real run."""

    def __init__(self, a):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        if self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(a.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        if self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy failed")

    def free(self):
        self.hip.hipFree(self.ptr)


def windows(fix, step0, steps, nwin, per_launch):
    """Median over nwin windows of `steps` timed steps (each window bracketed by stream
    synchronisations and a state read); returns (median seconds, all window seconds)."""
    ts = []
    for k in range(nwin):
        ts.append(timed_run(fix, step0 + k * steps, steps, per_launch))
    return float(np.median(ts)), ts


def cpu_info(orc):
    model, ncpu = orc.host_cpu()
    return {"cores": orc.threads(), "host_logical_cpus": ncpu, "cpu_model": model}


# ------------------------------------------------------------------ config 2 (headline)
def run_c2(mdt, args):
    log("config 2: building 256k-atom system")
    system = c2_system(mdt)
    n = system.n_local
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, system)
    # every step reads the positions from the engine's device buffer (packed at the
    # start of the step, inside its graph): the per-step position ingestion of an MD step
    xbuf = EngineBuffer(system.arrays()[0])
    ctx.attach_positions(0, xbuf.ptr.value)
    fix = mdt.Fix(ctx, C2_ARGS)
    fix.init()
    fix.run(0, args.warmup, per_launch=1)
    step = args.warmup
    # headline: one step at a time, what a caller that integrates between steps pays, as
    # one graph launch per step or as the step's kernels launched one by one (the faster
    # of the two is the headline; both are reported)
    dtg, tsg = windows(fix, step, args.steps, args.windows, 1)
    step += args.windows * args.steps
    log(f"c2, one graph launch per step: {dtg * 1e6 / args.steps:.1f} us/step (median of {args.windows}), "
        f"{n * args.steps / dtg:.3e} atom-steps/s")
    dtp, tsp = windows(fix, step, args.steps, args.windows, 0)
    step += args.windows * args.steps
    log(f"c2, plain kernel launches: {dtp * 1e6 / args.steps:.1f} us/step (median of {args.windows}), "
        f"{n * args.steps / dtp:.3e} atom-steps/s")
    plain = dtp < dtg
    dt, ts1 = (dtp, tsp) if plain else (dtg, tsg)
    value = n * args.steps / dt
    dt8, ts8 = windows(fix, step, args.steps, args.windows, 8)
    step += args.windows * args.steps
    log(f"c2, 8 steps per graph launch: {dt8 * 1e6 / args.steps:.1f} us/step, {n * args.steps / dt8:.3e} atom-steps/s")
    kms = fix.profile(step, args.profile_steps)
    log("kernel mean ms: " + json.dumps(kms))
    roof, pairs = steinhardt_roofline(fix, system, kms, fused_ld=True)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        x, off, idx, own = system.arrays()

        def one():
            env = orc.Env(x, off, idx, own, n, RC, 18)
            env.q_ave(6), env.q_ave(4), env.ld(6)
            -(env.dQ_dx(6, 1.0) + env.dQ_dx(4, 1.0))
        k, t = oracle_steps(one, args.cpu_budget)
        ci = cpu_info(orc)
        cpu = {"value": n * k / t, "unit": "atom-steps/s", "cores": ci["cores"], "kind": "port",
               "sample": f"{k} full step(s) of the same {n}-atom workload (selection, Q6, Q4, LQ6, exact "
                         f"Q6+Q4 forces) by the FP64 C++ oracle, OpenMP over atoms on {ci['cores']} threads "
                         f"({ci['cpu_model']}, {ci['host_logical_cpus']} logical CPUs), {t:.1f} s",
               "cpu_model": ci["cpu_model"], "host_logical_cpus": ci["host_logical_cpus"]}
    # whole-step roofline: sum over phases of max(bytes / HBM peak, FLOP / FP32 peak):
    # selection (HBM), forward q_lm, gradient, local-Q6 forward (FP32)
    t_roof = (roof["selection_bytes"] / (HBM_PEAK_GBS * 1e9) +
              (FLOP_FWD_PAIR + FLOP_GRAD_PAIR + FLOP_LD_FWD_PAIR) * pairs / (FP32_PEAK_TFLOPS * 1e12))
    step_roof = {"t_roof_us": t_roof * 1e6, "t_step_us": dt * 1e6 / args.steps, "frac": t_roof / (dt / args.steps),
                 "phases": "selection bytes / 8 TB/s + (261 + 903 + 28) FLOP x selected pairs / 157.3 TFLOP/s"}
    cfg = {"workload": "config 2: 256,000-atom FCC Ar half melted, Q6.AVE+Q4.AVE+LD-LQ6 (N=18, rc=3.9), "
                       "DIM v1+v2 [0,2]x2000 WT", "atoms": n, "selected_pairs": pairs, "kernels_ms": kms,
           "step": "positions read from the engine's device buffer, selection, Q6, Q4, LQ6.AVE, FP64 sums, "
                   "hill deposit and bias, chain rule, exact forces: one step at a time, "
                   + ("its kernels launched one by one" if plain else "one graph launch per step"),
           "windows_s": ts1, "step_roofline": step_roof,
           "one_graph_launch_per_step": {"value": n * args.steps / dtg, "ms_per_step": dtg * 1e3 / args.steps,
                                         "windows_s": tsg},
           "plain_kernel_launches": {"value": n * args.steps / dtp, "ms_per_step": dtp * 1e3 / args.steps,
                                     "windows_s": tsp},
           "eight_steps_per_graph_launch": {"value": n * args.steps / dt8, "ms_per_step": dt8 * 1e3 / args.steps,
                                            "windows_s": ts8},
           "neighbor_build": neighbor_build_ms(mdt, system)}
    xbuf.free()
    return line(value, args.steps, args.warmup, dt * 1e3 / args.steps, "f32 pair math, f64 sums/grid", cfg, roof, cpu)


def neighbor_build_ms(mdt, system, reps=5):
    """Ghost images and full lists of the same system rebuilt on the device (mdt_neighbor_build,
    outside the timed step: an MD loop rebuilds lists every few steps), mean of reps calls."""
    x = system.arrays()[0][: system.n_local]
    lo, hi = system.box()
    ctx = mdt.Context(0, 1)
    na, nz = ctx.build_neighbors(x, lo, hi, RG)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.build_neighbors(x, lo, hi, RG)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    ctx.close()
    return {"ms": ms, "atoms_with_ghosts": na, "list_entries": nz, "list_cutoff": RG}


# ------------------------------------------------------------------ config 1
def run_c1(mdt, args):
    s = c1_system(mdt)
    n = s.n_local
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, C1_ARGS)
    fix.init()
    fix.run(0, args.warmup)
    steps = 10000
    dt, mode, ts = timed_modes(fix, args.warmup, steps, (8, 0))
    value = n * steps / dt
    kms = fix.profile(args.warmup + 2 * steps, args.profile_steps)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        x, off, idx, own = s.arrays()

        def one():
            env = orc.Env(x, off, idx, own, n, RC, 18)
            env.q_ave(6)
            env.dQ_dx(6, 1.0)
        k, t = oracle_steps(one, min(args.cpu_budget, 10.0))
        ci = cpu_info(orc)
        cpu = {"value": n * k / t, "unit": "atom-steps/s", "cores": ci["cores"], "kind": "port",
               "sample": f"{k} steps of the 4,000-atom Q6 case (selection, Q6.AVE, exact forces), FP64 oracle, "
                         f"OpenMP over atoms on {ci['cores']} threads, {t:.1f} s"}
    cfg = {"workload": "config 1: 4,000-atom FCC Ar (sigma 0.1 A), Q6.AVE on [0,1]x1000, PACE 10, WT 0, "
                       "10,000 steps", "atoms": n, "kernels_ms": kms, "run_mode": mode_name(mode),
           "us_per_step_by_mode": {mode_name(m): t * 1e6 / steps for m, t in ts.items()}}
    return line(value, steps, args.warmup, dt * 1e3 / steps, "f32 pair math, f64 sums/grid", cfg, None, cpu,
                metric="atom-steps/s, config 1 (4,000 atoms, Q6.AVE)")


# ------------------------------------------------------------------ config 3
def run_c3(mdt, args):
    s = c3_system(mdt)
    n = s.n_local
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, C3_ARGS)
    fix.init()
    fix.run(0, 2)
    steps = max(5, args.steps // 100)
    dt, mode, ts = timed_modes(fix, 2, steps, (8, 0))
    value = n * steps / dt
    kms = fix.profile(2 + 2 * steps, 3)
    nk = 2056
    fl = FLOP_SK * n * nk
    tk = (kms.get("k_sk_amp", 0) + kms.get("k_sk_force", 0)) * 1e-3
    roof = {"kernel": "k_sk_amp+k_sk_force", "bound": "fp32", "achieved": fl / tk / 1e12, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": fl / tk / 1e12 / FP32_PEAK_TFLOPS, "traffic": None,
            "algorithmic": f"24 FLOP x {n} atoms x {nk} k per step"}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        x = s.arrays()[0][:n]
        rng = np.random.default_rng(7)
        budget = min(args.cpu_budget, 10.0)
        done_k, t0 = 0, time.perf_counter()
        while True:
            # chunks of 64 k vectors of the same magnitude range as the workload's
            m = rng.uniform(1.5, 3.5, size=64)
            v = rng.normal(size=(64, 3))
            ks = v / np.linalg.norm(v, axis=1)[:, None] * m[:, None]
            orc.sk(x, ks)
            orc.sk_grad(x, ks, 1.0)
            done_k += len(ks)
            t = time.perf_counter() - t0
            if t >= budget:
                break
        cores = cpu_info(orc)["cores"]
        # a full step is 2,056 vectors: scale the measured vectors per second to a step
        cpu = {"value": n / (t * nk / done_k), "unit": "atom-steps/s", "cores": cores, "kind": "port",
               "sample": f"S(k) and its gradient for {done_k} k vectors of the 500,000-atom slab (FP64 oracle, "
                         f"OpenMP, {cores} threads), {t:.1f} s, scaled by 2056/{done_k} to a full step"}
    cfg = {"workload": "config 3: 500,000-atom FCC/liquid slab, SK.AVE over 8 (111) + 2048 random k, "
                       "grid [0,50]x1000", "atoms": n, "k_vectors": nk, "cv_value": fix.cv("SK.AVE"),
           "kernels_ms": kms, "run_mode": mode_name(mode),
           "us_per_step_by_mode": {mode_name(m): t * 1e6 / steps for m, t in ts.items()}}
    return line(value, steps, 2, dt * 1e3 / steps, "f64 phase, f32 sincos, f64 sums", cfg, roof, cpu,
                metric="atom-steps/s, config 3 (500k atoms, S(k) over 2,056 k)")


# ------------------------------------------------------------------ config 4
def run_c4(mdt, args):
    systems = c4_systems(mdt)
    ntot = sum(s.n_local for s in systems)
    ctx = mdt.Context(0, len(systems))
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, C4_ARGS)
    fix.init()
    fix.run(0, args.warmup)
    steps = max(100, args.steps // 4)
    dt = timed_run(fix, args.warmup, steps)
    value = ntot * steps / dt
    kms = fix.profile(args.warmup + steps, args.profile_steps)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        s0 = systems[0]
        x, off, idx, own = s0.arrays()

        def one():
            env = orc.Env(x, off, idx, own, s0.n_local, RC, 18)
            env.q_ave(4), env.q_ave(6)
            -(env.dQ_dx(4, 1.0) + env.dQ_dx(6, 1.0))
        k, t = oracle_steps(one, min(args.cpu_budget, 10.0))
        ci = cpu_info(orc)
        cpu = {"value": s0.n_local * k / t, "unit": "atom-steps/s", "cores": ci["cores"], "kind": "port",
               "sample": f"{k} steps of walker 0 ({s0.n_local} atoms, Q4+Q6 and forces), FP64 oracle, OpenMP over "
                         f"atoms on {ci['cores']} threads, {t:.1f} s; the oracle runs walkers one after another"}
    cfg = {"workload": "config 4: 64 walkers, 2k-64k atoms each (FCC sigma 0.1 or liquid 0.021/A^3), "
                       "2-D grids 200x200 on (Q4.AVE, Q6.AVE), WT, one graph launch per step",
           "atoms_total": ntot, "walkers": len(systems), "kernels_ms": kms}
    return line(value, steps, args.warmup, dt * 1e3 / steps, "f32 pair math, f64 sums/grid", cfg, None, cpu,
                metric="atom-steps/s, config 4 (64 walkers, one graph launch per step)")


# ------------------------------------------------------------------ config 5
def run_c5(mdt, args):
    system = c2_system(mdt)
    n = system.n_local
    kt = 300.0 * KB
    base = C5_BASE
    res = {}
    steps = args.steps if args.plain else 2000
    nh0 = 19800
    hist = None
    for kind in args.c5_kinds.split(","):
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, system)
        extra = "Gaussian_Hill_type 2" if kind == "hills" else ("Gaussian_Hill_type 3 MLP_REFIT 500 "
                                                               "MLP_POINTS 1000000 MLP_HOLDOUT 65536")
        if kind == "mlp" and args.mlp_plain:
            extra += " MLP_GRAPH 0"
        fix = mdt.Fix(ctx, base + extra)
        fix.init()
        fix.post_force(0)
        if hist is None:
            s0 = np.array([fix.vector(k) for k in range(3)])
            # late-run state: 19,800 hills already deposited around the visited CV region
            hist = wt_hills(s0, nh0, kt)
        fix.add_hills(hist[0], hist[1], 0.01)
        t_fit = None
        if kind == "mlp":
            t0 = time.perf_counter()
            fix.mlp_refit()
            fix.sync()
            t_fit = time.perf_counter() - t0
        if args.plain:
            timed_plain(fix, 1, 20)
            dt = timed_plain(fix, 21, steps)
            r = {"run_mode": "host-driven post_force calls"}
        else:
            fix.run(1, 20)
            # the same steps (and refits) as 8 captured steps per graph launch and as plain launches
            dt, mode, ts = timed_modes(fix, 21, steps, (8, 0))
            r = {"run_mode": mode_name(mode), "us_per_step_by_mode": {mode_name(m): t * 1e6 / steps for m, t in ts.items()}}
        r.update({"atom_steps_per_s": n * steps / dt, "us_per_step": dt * 1e6 / steps,
                  "hills_end": int(fix.hills(3).shape[0])})
        if kind == "mlp":
            st = fix.mlp_status()
            r.update({"first_fit_s": t_fit, "heldout_max_dV_kT": st["err"] / kt, "network_in_use": st["ok"],
                      "fits": st["fits"], "fallbacks": st["fallbacks"]})
        res[kind] = r
        log(f"c5 {kind}: {r}")
        fix.close()
        ctx.close()
    cfg = {"workload": "config 5: 256,000-atom config-2 system, 3-D CV (Q4.AVE, Q6.AVE, S(k_111)), "
                       "sigma 0.01/0.01/0.5, h 0.005, WT 10; 19,800 hills pre-deposited, "
                       f"{steps} timed steps per run mode (+{steps // 10} hills each), learned bias refit every 500 steps on 1M points"
                       + (", host-driven plain launches" if args.plain else ""), "atoms": n, "results": res}
    if "mlp" in res:
        return line(res["mlp"]["atom_steps_per_s"], steps, 20, res["mlp"]["us_per_step"] * 1e-3,
                    "bf16 MFMA network (f32 accum), f64 labels", cfg, None, None,
                    metric="atom-steps/s, config 5 (learned bias; exact hills-sum result in config.results)")
    return line(res["hills"]["atom_steps_per_s"], steps, 20, res["hills"]["us_per_step"] * 1e-3, "f64", cfg,
                None, None, metric="atom-steps/s, config 5 (exact hills sum)")


def run_lq(mdt, args):
    """STEINH L 6 MEAN_SOLID (RATIONAL weights in r, FERMI f on LQ_6) on the 256k-atom system."""
    system = c2_system(mdt)
    n = system.n_local
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, system)
    fix = mdt.Fix(ctx, LQ_ARGS)
    fix.init()
    fix.run(0, 10)
    steps = max(20, args.steps // 20)
    dt, mode, ts = timed_modes(fix, 10, steps, (8, 0))
    value = n * steps / dt
    log(f"lq: {dt * 1e6 / steps:.1f} us/step ({mode_name(mode)}), {value:.3e} atom-steps/s")
    kms = fix.profile(10 + 2 * steps, 10)
    log("kernel mean ms: " + json.dumps(kms))
    x, off, idx, own = system.arrays()
    entries = float(off[-1])          # RATIONAL r_0 = 3 is above cutoff_eps on the whole 5.9 A list
    n_ghost = x.shape[0] - n
    fl = FLOP_LQ_PAIR * entries
    t = kms["k_lq_pair"] * 1e-3
    # k_lq_pair does 282 FLOP per entry over ~69 bytes: below the FP64 balance point (78.6 TFLOP/s
    # / 8 TB/s = 9.8 FLOP/B), so HBM bounds it.  Algorithmic bytes, each array once: per entry
    # (sigma, sigma') 16, list index 4, reverse ELL index 8, u 8, gradient written 24; per atom
    # offset 4, ELL base 8, S 8, position 32, H row 112, (<H,q>, <K,LQ>) 16; per ghost position 32.
    lq_bytes = LQ_PAIR_ENTRY_B * entries + LQ_PAIR_ATOM_B * n + 32.0 * n_ghost
    roof = {"kernel": "k_lq_pair", "bound": "hbm", "achieved": lq_bytes / t / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": lq_bytes / t / 1e9 / HBM_PEAK_GBS, "traffic": None,
            "algorithmic": f"{LQ_PAIR_ENTRY_B:.0f} B x {entries:.0f} list entries + {LQ_PAIR_ATOM_B:.0f} B x {n} atoms "
                           f"+ 32 B x {n_ghost} ghosts per step",
            "algorithmic_bytes": lq_bytes, "mean_ms": kms["k_lq_pair"],
            "fp64_tflops": fl / t / 1e12, "fp64_frac": fl / t / 1e12 / FP64_PEAK_TFLOPS}
    tp = os.path.join(ROOT, "profiles", "traffic_k_lq_pair_lq.json")
    if os.path.exists(tp):
        roof["traffic"] = json.load(open(tp)).get("bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        swr = dict(type="RATIONAL", r0=3.0, d0=0.0, n=6, m=12)
        swc = dict(type="FERMI", r0=0.35, alpha=10.0)

        def one():
            orc.lq_c(x, off, idx, n, own, 6, swr, swc, 1e-6, "MEAN_SOLID")
        k, tc = oracle_steps(one, args.cpu_budget)
        ci = cpu_info(orc)
        cpu = {"value": n * k / tc, "unit": "atom-steps/s", "cores": ci["cores"], "kind": "port",
               "sample": f"{k} full step(s) (CV and exact gradient) of the same {n}-atom workload by the FP64 C "
                         f"oracle, OpenMP over atoms on {ci['cores']} threads ({ci['cpu_model']}, "
                         f"{ci['host_logical_cpus']} logical CPUs), {tc:.1f} s",
               "cpu_model": ci["cpu_model"], "host_logical_cpus": ci["host_logical_cpus"]}
    cfg = {"workload": "STEINH L 6 MEAN_SOLID on the 256,000-atom half-melted FCC system (RATIONAL r_0 3.0 "
                       "weights on the 5.9 A list, FERMI r_0 0.35 alpha 10 on LQ_6), DIM [0,2]x2000 WT",
           "atoms": n, "list_entries": entries, "cv_value": fix.cv("LQ.MEAN_SOLID"), "kernels_ms": kms,
           "run_mode": mode_name(mode), "us_per_step_by_mode": {mode_name(m): t * 1e6 / steps for m, t in ts.items()},
           "flop_per_entry": {"k_lq_q": FLOP_LQ_Q, "k_lq_lq": FLOP_LQ_LQ, "k_lq_h": FLOP_LQ_H, "k_lq_pair": FLOP_LQ_PAIR},
           "path_fp64_tflops": (FLOP_LQ_Q + FLOP_LQ_LQ + FLOP_LQ_H + FLOP_LQ_PAIR) * entries / (dt / steps) / 1e12}
    return line(value, steps, 10, dt * 1e3 / steps, "f64", cfg, roof, cpu,
                metric="atom-steps/s, STEINH L 6 MEAN_SOLID CV+bias forces @256k atoms")


# ------------------------------------------------------------------ the reference's Q4 deck
REF_Q4_MS = 30.9375      # test/STEINH/Q4/log.lammps:72-84: fix metad, 9.2812 s / 300 steps, 32,000 atoms


def run_q4ref(mdt, args):
    """The reference's own STEINH Q4 deck (test/STEINH/Q4/in.test-steinhardt): 32,000 atoms of
    FCC a = 3.26 A (20^3 cells, jittered 0.02 A for its 10 K), STEINH Q 4 cutoff_r 6.0
    cutoff_Natoms 12 on lists of 6.0 + 2.0 skin (~250 entries: the near-list path, and the
    cap binds for every atom), PACE 1, WT, DIM [0,40] x 400, sigma 0.2, height 0.01; positions
    read from the engine buffer, one graph launch per step.  Its only on-path timing is the
    reference's log: 30.9 ms per step on an unstated NVIDIA GPU (a debug build), with fix nvt."""
    s = mdt.System().add_fcc(20, 20, 20, 3.26, 0.02, 0.02, 0.0, seed=10).build(8.0)
    n = s.n_local
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    xbuf = EngineBuffer(s.arrays()[0])
    ctx.attach_positions(0, xbuf.ptr.value)
    deck = ("GAUSSIAN 0.2 0.01 10.0 PACE 1 TEMP 10 HILLS_FILE NONE "
            "CAL NAME Q4 STEINH Q 4 all cutoff_Natoms 12 cutoff_r 6.0 "
            "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 40 400 v1 WT 1")
    fix = mdt.Fix(ctx, deck)
    fix.init()
    fix.run(0, args.warmup, per_launch=1)
    steps = max(100, args.steps // 4)
    dt, ts = windows(fix, args.warmup, steps, args.windows, 1)
    value = n * steps / dt
    log(f"q4ref: {dt * 1e6 / steps:.1f} us/step, {value:.3e} atom-steps/s "
        f"(reference log: {REF_Q4_MS:.1f} ms/step)")
    kms = fix.profile(args.warmup + args.windows * steps, args.profile_steps)
    log("kernel mean ms: " + json.dumps(kms))
    x, off, idx, own = s.arrays()
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        def one():
            env = orc.Env(x, off, idx, own, n, 6.0, 12)
            env.q_ave(4)
            env.dQ_dx(4, 1.0)
        k, t = oracle_steps(one, args.cpu_budget)
        ci = cpu_info(orc)
        cpu = {"value": n * k / t, "unit": "atom-steps/s", "cores": ci["cores"], "kind": "port",
               "sample": f"{k} steps (selection, Q4.AVE, exact forces) of the same 32,000-atom deck by the FP64 "
                         f"oracle, OpenMP over atoms on {ci['cores']} threads, {t:.1f} s"}
    cfg = {"workload": "the reference's STEINH Q4 deck: 32,000-atom FCC a=3.26 A, Q 4 N=12 rc=6.0, lists 8.0 A "
                       "(mean %.0f entries), PACE 1, WT" % (off[-1] / n),
           "atoms": n, "kernels_ms": kms, "windows_s": ts,
           "reference_log_ms_per_step": REF_Q4_MS,
           "reference_log_atom_steps_per_s": n / (REF_Q4_MS * 1e-3)}
    xbuf.free()
    return line(value, steps, args.warmup, dt * 1e3 / steps, "f32 pair math, f64 sums/grid", cfg, None, cpu,
                metric="atom-steps/s, the reference's STEINH Q4 deck (32,000 atoms)")


def run_copy(mdt, args):
    """Plain device copy of a known size: the calibration for HBM byte counters and the
    achievable copy bandwidth of the chip (reads 1 GiB, writes 1 GiB per launch)."""
    import torch
    n = 1 << 28
    src = torch.full((n,), 1.0, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    steps = max(5, args.steps // 100)
    t0 = time.perf_counter()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    nbytes = 2 * 4 * n
    return {"metric": "copy GB/s (read+write)", "value": nbytes / dt / 1e9, "unit": "GB/s", "n_gpus": 1,
            "steps": steps, "warmup": 3, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "device copy of 2^28 float32 (1 GiB read + 1 GiB written per launch)",
                       "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--windows", type=int, default=5, help="timed windows of --steps steps; the median is reported")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "lq", "q4ref", "copy"])
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mlp-plain", action="store_true",
                    help="c5: the learned bias trains with plain launches instead of a replayed captured epoch")
    ap.add_argument("--c5-kinds", default="hills,mlp",
                    help="c5: bias kinds to run, comma-separated, from hills (exact sum) and mlp (learned)")
    ap.add_argument("--plain", action="store_true",
                    help="c5: time --steps host-driven steps (plain launches) instead of 2,000 replayed graph steps")
    args = ap.parse_args()
    if any(k not in ("hills", "mlp") for k in args.c5_kinds.split(",")):
        ap.error("--c5-kinds takes hills and/or mlp")
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not built)", file=sys.stderr)
        sys.exit(2)
    import mdt
    if not os.path.exists(mdt.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    out = {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5, "lq": run_lq,
           "q4ref": run_q4ref, "copy": run_copy}[args.config](mdt, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
