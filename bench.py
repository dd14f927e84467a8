#!/usr/bin/env python3
"""Benchmark of the GODAR DEM step on one MI355X.

Workload (config 3 of the benchmark set): 512 independent floe fields, floe
counts log-uniform in [500, 8000] (config-0 generator per member, own seed
and wind direction), advanced together in one step pipeline in mixed
precision.  Reported beside it, on the same box:

* the plain FP64 single-simulation GPU path on the 1,000-floe field (config 0)
  and on the 262,144-floe field (config 1);
* the relative position error of the mixed-precision path after 10^4 steps
  of the 1,000-floe field against the fixture made by GODAR's own routines;
* the CPU baseline: GODAR's own physics routines compiled from the reference
  sources (oracle/_ref), or the C oracle where that is absent: one core, and
  one run per core on up to 16 cores at once (the aggregate is
  `cpu_baseline`).

`value` is whole-job floe-steps/s of the batched workload with all state
resident in HBM; the timed region is K steps bracketed by device syncs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(ROOT, "profiles", "step_kernels_batched_512fields.md")
STAGE_KERNEL = {"candidates": "k_cand", "contacts": "k_contact", "bonds": "k_bond",
                "walls": "k_walls", "repair": "k_repair",
                "fixup": "k_fixup", "integrate": "k_final"}


def profile_traffic(stage: str, precision: str):
    """HBM bytes per launch (read + write, counter-scaled) of a stage's kernel
    from the committed rocprofv3 summary, or None.  The table gives MB
    (10^6 bytes) or MiB per launch, as its header says."""
    name = STAGE_KERNEL.get(stage)
    if not name or not os.path.exists(PROFILE):
        return None
    tmpl = "<true>" if precision == "mixed" else "<false>"
    unit = 1e6
    for line in open(PROFILE):
        if line.startswith("| kernel") and "MiB" in line:
            unit = float(1 << 20)
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) != 5 or not cols[0].startswith("`godar::"):
            continue
        kname = cols[0].strip("`")[len("godar::"):]
        if kname in (name + tmpl, name):
            try:
                return (float(cols[3]) + float(cols[4])) * unit
            except ValueError:
                return None
    return None


# the headline metric: batched floe-steps per second (with the plain FP64 and
# served figures beside it)
METRIC = "floe-steps/s batched vs plain FP64 on 1 GPU at stated pos err; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)





def build_batch(members: int, seed: int):
    from godar_amd import workloads as W
    t = time.time()
    batch = W.config3(seed=seed, members=members)
    log(f"[bench] generated {len(batch)} members, {sum(f.n for f, _ in batch)} floes "
        f"in {time.time() - t:.1f}s")
    return batch


def pair_totals(ctx, mids):
    t = ctx.totals()
    return t["candidates"], t["contacts"], t["bonds"], t


STAGE_INDEX = ["bin", "sort", "cells", "candidates", "contacts", "bonds", "unused", "walls",
               "repair", "fixup", "stress", "integrate"]


def run_batched(args, batch):
    """The headline leg: W warm-up steps and the autotuner, then (untimed)
    on to the middle of the configured run, K timed steps there, and K more
    with HIP events around every stage."""
    from godar_amd.sim import Context
    from godar_amd.workloads import algorithmic_bytes
    nfloes = sum(f.n for f, _ in batch)
    ctx = Context(nfloes, len(batch), args.precision)
    mids = [ctx.add_field(p, f) for f, p in batch]
    if args.no_graphs:
        ctx.use_graphs(False)
    ctx.step(args.warmup)
    if args.block:
        ctx.set_block(args.block)
        blocks = ctx.stage_blocks()
    else:
        blocks = ctx.autotune(2)      # real steps: part of the warm-up
    ctx.sync()
    search = ctx.search()
    graphs = ctx.graphs_on()
    launch = "graph replay" if graphs else "one-by-one launches"
    # the timed window starts at step `window_start` of the run (the steps
    # before it are untimed warm-up); later steps of the configured run hold
    # more contacts than the first ones
    t_next = ctx.tstep(mids[0])
    if args.window_start > t_next:
        ctx.step(args.window_start - t_next)
    ctx.sync()
    w0 = ctx.tstep(mids[0])
    # timed region: steps in the launch mode the autotuner chose (graph
    # replay or one-by-one launches), nothing else on the GPU
    t0 = time.perf_counter()
    ctx.step_async(args.steps)
    ctx.sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    ncand, ncont, nbond, tot = pair_totals(ctx, mids)
    # the same number of steps again with HIP events around every stage
    # (kernels launched one by one) for the per-kernel durations
    ctx.profile(True)
    ctx.step_async(args.steps)
    ctx.sync()
    stages = ctx.stage_times()
    ctx.profile(False)
    ctx.close()
    return dict(nfloes=nfloes, members=len(batch), seconds=dt, stages=stages, blocks=blocks,
                launch=launch, graphs=graphs, search=search, window=(w0, w0 + args.steps - 1),
                ncand=ncand, ncont=ncont, nbond=nbond, totals=tot,
                alg_bytes_step=algorithmic_bytes(nfloes, ncand, ncont, nbond))


def run_full(batch, blocks, graphs, search, precision, total=5000, chunk=500):
    """The configured run from step 1 on a fresh ensemble, with the block
    sizes, search shape and launch mode the autotuner chose: wall time of
    every `chunk` steps (device syncs around each) and the pairs per floe at
    its end."""
    from godar_amd.sim import Context
    nfloes = sum(f.n for f, _ in batch)
    ctx = Context(nfloes, len(batch), precision)
    for f, p in batch:
        ctx.add_field(p, f)
    for k, name in enumerate(STAGE_INDEX):
        if name in blocks and name != "unused":
            try:
                ctx.set_stage_block(k, blocks[name])
            except Exception:
                pass
    ctx.set_search(**search)
    ctx.use_graphs(graphs)
    ctx.sync()
    phases = []
    whole = 0.0
    for a in range(1, total + 1, chunk):
        t0 = time.perf_counter()
        ctx.step_async(chunk)
        ctx.sync()
        el = time.perf_counter() - t0
        whole += el
        tot = ctx.totals()
        phases.append(dict(steps=f"{a}-{a + chunk - 1}", ms_per_step=el / chunk * 1e3,
                           floe_steps_per_s=nfloes * chunk / el,
                           candidates_per_floe=tot["candidates"] / nfloes,
                           contacts_per_floe=tot["contacts"] / nfloes,
                           bonds_per_floe=tot["bonds"] / nfloes))
        log(f"[bench] full run steps {a}-{a + chunk - 1}: {el / chunk * 1e3:.3f} ms/step, "
            f"{tot['contacts']} contacts")
    ctx.close()
    return dict(steps=total, floes=nfloes, seconds=whole, ms_per_step=whole / total * 1e3,
                floe_steps_per_s=nfloes * total / whole, launch_mode=("graph replay" if graphs
                                                                      else "one-by-one launches"),
                phases=phases)


def run_serving(nreq: int, seed: int = 4, hbm_gib: float = 16.0, chunk: int = 25):
    """Served requests (Poisson arrivals at 20/s, 1,000-50,000 floes,
    500-5,000 steps, mixed precision) through the continuous-batching loop."""
    from godar_amd import workloads as W
    from godar_amd.serve import Request, Server
    t = time.time()
    reqs = [Request(k, f, p, s, a) for k, (f, p, s, a) in enumerate(W.config4_requests(nreq, seed))]
    log(f"[bench] generated {nreq} served requests in {time.time() - t:.1f}s")
    srv = Server(hbm_bytes=int(hbm_gib * (1 << 30)), max_members=256, precision="mixed",
                 chunk=chunk)
    # warm the context (graph capture, first launches) on a throwaway request
    warm = Request(-1, reqs[0].field, reqs[0].params, 3, 0.0)
    srv.run([warm])
    out = srv.run(reqs)
    srv.close()
    return dict(requests=out["requests"], arrival_rate_per_s=20.0, precision="mixed",
                hbm_budget_gib=hbm_gib, capacity_floes=out["capacity_floes"],
                floe_steps=out["floe_steps"], makespan_s=out["makespan_s"],
                floe_steps_per_s=out["floe_steps_per_s"],
                latency_p50_s=out["latency_p50_s"], latency_p99_s=out["latency_p99_s"],
                peak_members=out["peak_members"], peak_floes=out["peak_floes"], chunk=chunk)


def run_served_mix(nreq: int, steps: int, seed: int = 4):
    """The served requests' fields stepped together as one batch (every
    request admitted at once, mixed precision): 5 warm-up steps, then `steps`
    timed steps and the per-stage times of as many more."""
    from godar_amd import workloads as W
    from godar_amd.sim import Context
    reqs = W.config4_requests(nreq, seed)
    n = sum(f.n for f, _, _, _ in reqs)
    ctx = Context(n, len(reqs), "mixed")
    for f, p, _, _ in reqs:
        ctx.add_field(p, f)
    ctx.step(5)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    ms = (time.perf_counter() - t0) / steps * 1e3
    ctx.profile(True)
    ctx.step(steps)
    ctx.sync()
    st = ctx.stage_times()
    ctx.profile(False)
    out = {"workload": f"{nreq} served requests at once", "floes": n, "steps": steps,
           "ms_per_step": ms,
           # per step: the build stages launch only on steps with the build kernels
           "stages_ms_per_step": {k: v[0] / steps for k, v in st.items() if v[1]},
           "pairs": ctx.totals()}
    ctx.close()
    return out


def run_plain(which: str, steps: int, warmup: int):
    from godar_amd import workloads as W
    from godar_amd.sim import Context
    f, p = W.config0() if which == "config0" else W.config1()
    ctx = Context(f.n, 1, "fp64", outer=128 if which == "config1" else 0,
                  inner=128 if which == "config1" else 0)
    ctx.add_field(p, f)
    ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step_async(steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.close()
    return f.n * steps / dt, dt / steps * 1e3


def pos_error_1e4(precision="mixed"):
    """Relative position error after 10^4 steps of the 1,000-floe field against
    the fixture made by GODAR's own routines: (max, median, envelope max,
    envelope median), the envelope being the reference's own deviation when
    every initial centre moves by one ulp."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from fixtures import load, load_long
    from godar_amd.sim import Context
    f, p, _ = load("config0_seed0.npz")
    L = load_long()
    ctx = Context(f.n, 1, precision)
    m = ctx.add_field(p, f)
    ctx.step(10000)
    a = ctx.floes(m)
    ctx.close()
    xy = np.concatenate([a["x"], a["y"]])
    ref = np.concatenate([L["x_10000"], L["y_10000"]])
    rel = np.abs(xy - ref) / np.abs(ref)
    env_max = max(float(L["ulp_env_max_10000_0"]), float(L["ulp_env_max_10000_1"]))
    env_med = max(float(L["ulp_env_median_10000_0"]), float(L["ulp_env_median_10000_1"]))
    uv = np.concatenate([a["u"], a["v"]])
    uv_ref = np.concatenate([L["u_10000"], L["v_10000"]])
    vol = lambda r, h: float(np.sum(np.pi * r * r * h))        # ridging conserves each floe's pi r^2 h
    v0 = vol(f.r, f.h)
    return dict(max=float(rel.max()), median=float(np.median(rel)),
                rel_l2=float(np.linalg.norm(xy - ref) / np.linalg.norm(ref)),
                vel_rel_l2=float(np.linalg.norm(uv - uv_ref) / np.linalg.norm(uv_ref)),
                vel_max_abs=float(np.abs(uv - uv_ref).max()),
                volume_drift=(vol(a["r"], a["h"]) - v0) / v0,
                volume_drift_reference=(vol(L["r_10000"], L["h_10000"]) - v0) / v0,
                ulp_envelope_max=env_max, ulp_envelope_median=env_med)


def cpu_baseline(target_seconds: float):
    """GODAR's own routines (oracle/_ref) on the 1,000-floe field, one core
    (a plain dict, so that worker processes can return it)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from godar_amd import workloads as W
    f, p = W.config0()
    import ref_harness as rh
    cfg = "n1000fast" if rh.available("n1000fast") else "n1000"
    if rh.available(cfg):
        H = rh.RefHarness(cfg)
        H.setup(p, f, bondmask=rh.dense_bondmask(f.n, p.bond_seed, p.bond_frac), logcap=1)
        step = H.step
        kind = "reference"
        flags = ("-O3 -funroll-loops -ffast-math, the reference's Release flags"
                 if cfg == "n1000fast" else "-O2 -ffp-contract=off")
        what = (f"GODAR physics routines compiled from the reference sources (amdflang {flags}), "
                "stepper control flow restated, cell-list radius search")
    else:
        from oracle_py import OracleSim
        O = OracleSim(p, f)
        step = lambda t: O.step()
        kind = "port"
        what = "C oracle (restatement of the reference step)"
    t0 = time.perf_counter()
    t = 0
    while True:
        t += 1
        step(t)
        el = time.perf_counter() - t0
        if el >= target_seconds:
            break
    return dict(value=f.n * t / el, unit="floe-steps/s", cores=1, kind=kind, steps=t, seconds=el,
                sample=f"config 0 (1,000 floes, 100 km box), first {t} steps, {el:.1f} s; {what}")


def host_info() -> dict:
    """What the host offers: logical CPUs, the CPUs this process may run on,
    and the processor model."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return dict(logical_cpus=os.cpu_count(), cpus_allowed=allowed, model=model)


def _cpu_worker(target_seconds):
    return cpu_baseline(target_seconds)


def cpu_baseline_all_cores(target_seconds: float, workers: int):
    """The same run in `workers` processes at once (the batched workload's
    members are independent, so a CPU runs one per core): aggregate
    floe-steps/s.  Called before the GPU is touched (worker processes are
    forked)."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        runs = pool.map(_cpu_worker, [target_seconds] * workers)
    one = runs[0]
    total = sum(r["value"] for r in runs)
    return dict(value=total, unit="floe-steps/s", cores=workers, kind=one["kind"],
                sample=(f"{workers} processes, each config 0 (1,000 floes, 100 km box) for "
                        f"{target_seconds:.0f} s ({min(r['steps'] for r in runs)}-"
                        f"{max(r['steps'] for r in runs)} steps); "
                        + one["sample"].split("; ", 1)[1]),
                per_core=[r["value"] for r in runs])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--members", type=int, default=512)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--window-start", type=int, default=2501,
                    help="step of the configured 5,000-step run at which the timed window starts")
    ap.add_argument("--full-steps", type=int, default=5000,
                    help="steps of the full-run leg (0: skip)")
    ap.add_argument("--quick", action="store_true", help="batched run only (for profilers)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--serve-requests", type=int, default=100,
                    help="served requests of the continuous-batching leg (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates rocprof byte counters)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch kernels one by one (no graph replay) in the batched leg")
    ap.add_argument("--served-mix", action="store_true",
                    help="only step all served requests at once (for profilers)")
    args = ap.parse_args()
    if args.served_mix:
        print(json.dumps(run_served_mix(args.serve_requests, args.steps)))
        return
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": a.numel() * 8,
                          "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is documented, not built)", file=sys.stderr)
        sys.exit(2)
    cb = cb1 = None
    if not args.quick:
        # CPU baselines first: the worker processes are forked before any GPU use.
        # A GPU box allots 16 host cores to its GPU: one process per core, at
        # most 16, whatever the machine has in total.
        host = host_info()
        workers = max(1, min(16, host["cpus_allowed"]))
        cb1 = cpu_baseline(args.cpu_seconds)
        cb = cpu_baseline_all_cores(args.cpu_seconds, workers) if workers > 1 else cb1
    batch = build_batch(args.members, args.seed)
    B = run_batched(args, batch)
    value = B["nfloes"] * args.steps / B["seconds"]
    ms_step = B["seconds"] / args.steps * 1e3
    # dominant stage by measured time; roofline on its algorithmic bytes
    stages = {k: v for k, v in B["stages"].items() if v[1] > 0}
    # per step (the build stages launch only on steps with the build kernels)
    # and per launch (the kernel's own duration, for the roofline)
    dom = max(stages, key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    nf, nc, nk, nb = B["nfloes"], B["ncand"], B["ncont"], B["nbond"]
    kbar, kc, kb = nc / nf, nk / nf, nb / nf
    qc, qb = B["totals"]["queued_contact"], B["totals"]["queued_bond"]
    # algorithmic bytes per launch of each stage (partner records are re-reads
    # of other floes' rows and count once, in the floe's own term):
    #  shelter: x, y, r, hfa, hfw + wind read, 2 minima written, history and
    #    bond counts read, class written; per candidate pair one list entry
    #    read on each side, flag and override words written
    #  contacts: queued floe's 11 words read, 5 accumulators + count written;
    #    per contact two list entries read, history (16 B) read and written
    #  bonds: same per queued floe; per bond two list entries read, the
    #    88-byte record read and written
    #  integrate: the floe state (12 words) read, the 6 evolving words
    #    written, 2 sheltering minima read; 5 force accumulators per contact
    #    floe; per bond its 10 terms read by each floe (the bond gather)
    per_stage_bytes = {
        "shelter": nf * (7 * 8 + 2 * 8 + 2 * 4 + 1) + nc * (2 * 8 + 4 + 4),
        "contacts": qc * (11 * 8 + 5 * 8 + 4) + nk * (2 * 8 + 32),
        "bonds": qb * (11 * 8 + 5 * 8 + 4) + nb * (2 * 8 + 2 * 88),
        "candidates": nf * (4 * 8 + 8) + 24.0 * nc,
        "integrate": nf * (12 * 8 + 6 * 8 + 2 * 8) + qc * 5 * 8 + nb * 2 * 10 * 8,
        "sort": nf * 8.0 * 2 * 3,
        #  repair: per event re-evaluated, both floes' current properties and
        #    rows (2 x (48 + 32 + 56) B) and the pair state (40 B history, 88 B
        #    bond record) read, its 264-byte result written
        "repair": B["totals"]["repaired_events"] * (2 * (48 + 32 + 56) + 40 + 88 + 264),
    }
    # the roofline is the dominant kernel's; a stage without a byte model
    # (fix-up, walls, build) passes the place to the next one
    modelled = {k: v for k, v in stages.items() if k in per_stage_bytes}
    top = dom
    dom = max(modelled, key=lambda k: modelled[k][0] / modelled[k][1])
    dom_ms = stages[dom][0] / stages[dom][1]
    alg = per_stage_bytes[dom]
    achieved = alg / (dom_ms * 1e-3) / 1e9
    traffic = profile_traffic(dom, args.precision)
    step_gbs = B["alg_bytes_step"] / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "floe-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"config 3: {B['members']} independent floe fields, "
                         f"{nf} floes (log-uniform 500-8000 per field), {args.precision} precision, "
                         "one step pipeline"),
            "window": (f"steps {B['window'][0]}-{B['window'][1]} of the configured 5,000-step run "
                       "(earlier steps: untimed warm-up and autotuning)"),
            "members": B["members"], "floes": nf, "precision": args.precision,
            "candidates_per_floe": kbar, "contacts_per_floe": kc, "bonds_per_floe": kb,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": ("profiles/step_kernels_batched_512fields.md (rocprofv3 --pmc "
                               "FETCH_SIZE x2.000, WRITE_SIZE x1.000, bytes per launch)"
                               if traffic is not None else
                               "null: rocprofv3 --pmc ends in a fault inside the profiler on this "
                               "workload (profiles/step_kernels_batched_512fields.md)"),
            "kernel_avg_ms": dom_ms, "algorithmic_bytes": alg, "dominant_stage": top,
            "step_algorithmic_GBs": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS,
        },
        "stages_ms_per_step": {k: v[0] / args.steps for k, v in stages.items()},
        "pair_totals_last_step": B["totals"],
        "block_sizes": B["blocks"],
        "search_shape": B["search"],
        "launch_mode": B["launch"],
    }
    if not args.quick and args.full_steps > 0:
        out["full_run"] = run_full(batch, B["blocks"], B["graphs"], B["search"], args.precision,
                                   total=args.full_steps)
    if not args.quick:
        p0, ms0 = run_plain("config0", 500, 20)
        p1, ms1 = run_plain("config1", 20, 3)
        out["plain_fp64"] = {"config0_floe_steps_s": p0, "config0_ms_per_step": ms0,
                             "config1_floe_steps_s": p1, "config1_ms_per_step": ms1}
        out["batched_vs_plain_config0"] = value / p0
        out["batched_vs_plain_config1"] = value / p1
        out["pos_err_1e4"] = {"mixed": pos_error_1e4("mixed"), "fp64": pos_error_1e4("fp64")}
        if args.serve_requests > 0:
            out["serving"] = run_serving(args.serve_requests)
        out["cpu_baseline"] = dict(cb, host=host, share=(
            f"{workers} processes: the host cores a GPU box allots to its one GPU (16) "
            f"of {host['cpus_allowed']} this process may use ({host['logical_cpus']} logical CPUs)"))
        out["cpu_baseline_one_core"] = {k: cb1[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["vs_cpu_baseline"] = value / cb["value"]
        out["vs_cpu_baseline_one_core"] = value / cb1["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
