"""Benchmark of the HydroCore time step on one MI355X.

Workload (the configuration the headline metric is quoted on): the clumpy
self-gravitating medium on a 512^3 mesh -- lognormal density (mean n = 1
cm^-3, sigma_ln rho = 1) from a seeded k^-4 Gaussian random field, RMS Mach 2
velocities, T = 1e4 K, a central 1e4 Msun point mass softened over 2 cells,
periodic boundaries -- advanced with HLLC + van Leer, FP32 sweeps with FP64
energy recovery and accumulation, artificial viscosity, exact tabulated
cooling (Z1.0.dat) and point-mass gravity.  A step is one iteration of the
reference's main loop (source/mainHD.f90:74-99): two fused kernels.

Prints one JSON line: zone-updates/s over K timed steps (inputs resident in
HBM), the HBM roofline of the dominant kernel (stage2), the CPU baseline (the
FP64 C oracle with OpenMP on a bounded sample of the same workload), the L1
error of the FP32 path against the FP64 path on the same problem, and the
cooling paths compared on the same mesh: the reference's one-step formula
fused into stage 2 (the headline value), the learned surrogate with its
per-cell fallback (trained in this run on the GPU), and the converged
integrator (the surrogate's target and fallback), on the workload as given,
at T = 3e5 K, and at T = 3e5 K with every density times 100 (strongly
cooling: dt/t_cool of order 1-10, hundreds of integrator substeps per cell).
A last leg runs the same workload with self-gravity (the multigrid Poisson
solve of MultiGrid, multipole boundary values every 10 steps) and reports
its V-cycles and time per step.
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

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
BYTES_STAGE1 = 40           # per zone: read U^n, write U^{1/2} (5 fields x 4 B each)
BYTES_STAGE2 = 60           # per zone: read U^{1/2} and U^n, write U^{n+1}
BYTES_STEP = BYTES_STAGE1 + BYTES_STAGE2


TRAFFIC_REPORT = os.path.join(ROOT, "profiles", "traffic_clumpy512.md")
SG_TRAFFIC_REPORT = os.path.join(ROOT, "profiles", "traffic_selfgravity512.md")
BYTES_RELAX0 = 20           # per cell and red-black iteration: read PHI and rho, write PHI


def sg_traffic():
    """Measured HBM bytes of one finest-level relaxation launch (calibrated
    rocprof counters), or None."""
    if not os.path.exists(SG_TRAFFIC_REPORT):
        return None
    from tlaloque_amd.profiling import read_traffic
    k = read_traffic(SG_TRAFFIC_REPORT)["kernels"].get("k_sg_rbz")
    return k["bytes"] if k else None
COPY_PROBE_BYTES = 2 << 30


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_run(p, u0, table, warmup, steps, theta=None):
    """Advance a fresh context over `steps` timed steps; returns
    (zone-updates/s, ms per step, per-kernel ms, (queued cells, fallbacks) of
    the last step, bad cells)."""
    from tlaloque_amd.hydro import HydroCore
    with HydroCore(p, cooling_table=table) as h:
        if theta is not None:
            h.surrogate_params(theta)
        h.set_state(u0)
        h.advance(warmup)
        h.sync()
        h.kernel_timing(True)
        h.timer_start()
        h.advance(steps)
        ms = h.timer_stop()
        ks = h.kernel_stats()
        h.kernel_timing(False)
        counts = h.cooling_counts()
        bad = h.check_state()
    kms = {k: v[0] / max(v[1], 1) for k, v in ks.items() if v[1]}
    return p.zones * steps / (ms / 1e3), ms / steps, kms, counts, bad


def heated(p, u0, temp, dens=1.0):
    """The same state at temperature `temp` (every cell of the clumpy set-up
    starts at one temperature) with all densities scaled by `dens` (the
    velocities, hence the hydro time step, unchanged)."""
    from tlaloque_amd import params as P
    u = u0.copy()
    rho = u[..., 0]
    ek = 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / rho
    eth = u[..., 4] - ek
    t0 = float(np.median((p.gamma - 1.0) * eth * p.mu0 * P.AMU / (rho * P.KB)))
    u[..., 4] = (ek + eth * (temp / t0)) * dens
    u[..., :4] *= dens
    return u


def cooling_modes(p, u0, table, warmup, steps):
    """The three cooling paths on the workload and on its T = 3e5 K variant."""
    from tlaloque_amd import params as P
    from tlaloque_amd import surrogate
    t0 = time.perf_counter()
    theta, rep = surrogate.train()
    log(f"bench: surrogate trained in {time.perf_counter() - t0:.1f}s: {rep}")
    out = {"surrogate_training": {"samples": rep.samples, "steps": rep.steps,
                                  "batch": rep.batch, "seconds": rep.seconds,
                                  "heldout_rms_dex": rep.heldout_rms,
                                  "heldout_rel_median": rep.heldout_rel_median,
                                  "heldout_rel_p90": rep.heldout_rel_p90}}
    variants = (("T=1e4K", lambda: u0), ("T=3e5K", lambda: heated(p, u0, 3e5)),
                ("T=3e5K,n*100", lambda: heated(p, u0, 3e5, 100.0)))
    for label, make in variants:
        u = make()
        res = {}
        for name, mode in (("reference_one_step", P.COOL_EXACT),
                           ("surrogate", P.COOL_SURROGATE),
                           ("converged_integrator", P.COOL_SUBCYCLED)):
            if label == "T=1e4K" and mode == P.COOL_EXACT:
                continue  # the headline run
            zps, ms, kms, (nq, nf), bad = timed_run(p.replace(cooling=mode), u, table, warmup,
                                                     steps,
                                                     theta if mode == P.COOL_SURROGATE else None)
            res[name] = {"zone_updates_per_s": zps, "ms_per_step": ms,
                         "cooling_pass_ms": kms.get("cooling"), "nonfinite_cells": bad}
            if mode != P.COOL_EXACT:
                res[name]["queued_fraction"] = nq / p.zones
            if mode == P.COOL_SURROGATE:
                res[name]["fallback_fraction"] = nf / nq if nq else 0.0
            log(f"bench: {label} {name}: {zps:.4g} zone-updates/s, {ms:.3f} ms/step, "
                f"kernels {kms}, queued {nq}, fallbacks {nf}")
        if "surrogate" in res and "converged_integrator" in res:
            res["surrogate_speedup_vs_converged"] = (res["surrogate"]["zone_updates_per_s"] /
                                                     res["converged_integrator"]["zone_updates_per_s"])
            cs, ci = res["surrogate"]["cooling_pass_ms"], res["converged_integrator"]["cooling_pass_ms"]
            if cs and ci:
                res["cooling_pass_speedup"] = ci / cs
        out[label] = res
        del u
    return out


def self_gravity_leg(p, u0, table, warmup, steps):
    """The workload with self-gravity (TLQ_SG_MULTIPOLE): zone-updates/s of
    whole steps, multigrid ms and V-cycles per step."""
    from tlaloque_amd import params as P
    from tlaloque_amd.hydro import HydroCore
    q = p.replace(self_gravity=P.SG_MULTIPOLE)
    with HydroCore(q, cooling_table=table) as h:
        t0 = time.perf_counter()
        h.set_state(u0)
        lv, c0, _, _ = h.gravity_info()
        log(f"bench: self-gravity start-up solve: {c0} V-cycles, {time.perf_counter() - t0:.2f}s")
        h.advance(warmup)
        h.sync()
        h.kernel_timing(True)
        h.timer_start()
        cycles = []
        for _ in range(steps):
            h.advance(1)
            cycles.append(h.gravity_info()[1])
        ms = h.timer_stop()
        ks = h.kernel_stats()
        h.kernel_timing(False)
        bad = h.check_state()
    mg_ms = ks["multigrid"][0] / max(ks["multigrid"][1], 1)
    r0_ms = ks["mg_relax0"][0] / max(ks["mg_relax0"][1], 1)
    # one red-black iteration on the finest level: read PHI (8 B) and the FP32
    # density (4 B), write PHI (8 B) per cell
    ach = BYTES_RELAX0 * q.zones / (r0_ms / 1e3) / 1e9
    out = {"zone_updates_per_s": q.zones * steps / (ms / 1e3), "ms_per_step": ms / steps,
           "multigrid_ms_per_step": mg_ms, "levels": lv,
           "vcycles_per_step": float(np.mean(cycles)), "vcycles_startup": c0,
           "ms_per_vcycle": mg_ms / max(float(np.mean(cycles)), 1e-9),
           "boundary": "multipole every 10 steps", "nonfinite_cells": bad,
           "roofline": {"bound": "hbm", "kernel": "k_sg_rbz (finest level)", "achieved": ach,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "bytes_per_zone": BYTES_RELAX0, "avg_launch_ms": r0_ms,
                        "launches": ks["mg_relax0"][1], "traffic": sg_traffic()}}
    log(f"bench: self-gravity: {out}")
    return out


def host_cpu() -> dict:
    """The host's CPU model and logical CPU count (Linux /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(n: int, steps: int, threads: int):
    """The FP64 C oracle on the same workload at n^3 (bounded sample)."""
    import oracle
    from tlaloque_amd import params as P
    from tlaloque_amd import problems
    p, u0 = problems.clumpy(n=n, precision=P.PREC_FP64)
    table = P.load_cooling_table(os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat"))
    o = oracle.Oracle(p, table=table)
    o.set_cons(u0)
    oracle.set_threads(threads)
    o.advance()  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        o.advance()
    el = time.perf_counter() - t0
    return n ** 3 * steps / el, el


def l1_fp32_vs_fp64(n: int, steps: int):
    """L1 (sum|d|/sum|ref|) of rho, p and |v| of the FP32 path against the FP64
    path after `steps` steps of the same clumpy problem at n^3."""
    from tlaloque_amd import params as P
    from tlaloque_amd import problems
    from tlaloque_amd.hydro import HydroCore
    table = P.load_cooling_table(os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat"))
    out = {}
    for prec in (P.PREC_FP64, P.PREC_FP32):
        p, u0 = problems.clumpy(n=n, precision=prec)
        with HydroCore(p, cooling_table=table) as h:
            h.set_state(u0)
            h.advance(steps)
            u = h.download("U")[2:-2, 2:-2, 2:-2]
        rho = u[..., 0]
        v = u[..., 1:4] / rho[..., None]
        pr = (p.gamma - 1.0) * (u[..., 4] - 0.5 * rho * (v ** 2).sum(-1))
        out[prec] = (rho, pr, np.sqrt((v ** 2).sum(-1)))
    l1 = lambda a, b: float(np.abs(a - b).sum() / np.abs(b).sum())  # noqa: E731
    r64, r32 = out[P.PREC_FP64], out[P.PREC_FP32]
    return {"rho": l1(r32[0], r64[0]), "p": l1(r32[1], r64[1]), "v": l1(r32[2], r64[2]),
            "mesh": n, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-l1", action="store_true")
    ap.add_argument("--no-modes", action="store_true", help="skip the cooling-path comparison")
    ap.add_argument("--no-gravity", action="store_true", help="skip the self-gravity leg")
    ap.add_argument("--gravity-only", action="store_true",
                    help="run only the self-gravity leg and print its JSON (for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)

    from tlaloque_amd import params as P
    from tlaloque_amd import problems
    from tlaloque_amd.hydro import HydroCore

    n = args.n
    t0 = time.perf_counter()
    p, u0 = problems.clumpy(n=n, precision=P.PREC_FP32)
    log(f"bench: clumpy {n}^3 initial state built in {time.perf_counter() - t0:.1f}s")
    table = P.load_cooling_table(os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat"))
    if args.gravity_only:
        print(json.dumps({"self_gravity": self_gravity_leg(p, u0, table, args.warmup, args.steps),
                          "config": {"workload": f"clumpy self-gravitating medium {n}^3 with "
                                                 "self-gravity (multipole boundaries)"}}),
              flush=True)
        return
    h = HydroCore(p, cooling_table=table)
    h.set_state(u0)
    h.advance(args.warmup)
    h.sync()
    log(f"bench: {args.warmup} warm-up steps done")
    h.kernel_timing(True)
    h.timer_start()
    h.advance(args.steps)
    ms = h.timer_stop()
    ks = h.kernel_stats()
    h.kernel_timing(False)
    bad = h.check_state()
    dt, t, it = h.time()
    copy_ms = h.copy_probe(COPY_PROBE_BYTES, 10)
    h.close()
    zones = n ** 3
    value = zones * args.steps / (ms / 1e3)
    s2_ms = ks["stage2"][0] / max(ks["stage2"][1], 1)
    s1_ms = ks["stage1"][0] / max(ks["stage1"][1], 1)
    achieved = BYTES_STAGE2 * zones / (s2_ms / 1e3) / 1e9
    copy_gbs = 2 * COPY_PROBE_BYTES / (copy_ms / 1e3) / 1e9
    log(f"bench: {args.steps} steps in {ms:.1f} ms; stage1 {s1_ms:.3f} ms, stage2 {s2_ms:.3f} ms; "
        f"bad cells {bad}; t={t:.4e} it={it}; copy probe {copy_gbs:.0f} GB/s")
    traffic, traffic_note = None, None
    if n == 512 and os.path.exists(TRAFFIC_REPORT):
        from tlaloque_amd.profiling import read_traffic
        tr = read_traffic(TRAFFIC_REPORT)
        if tr.get("zones") == zones and "k_stage2" in tr["kernels"]:
            traffic = tr["kernels"]["k_stage2"]["bytes"]
            traffic_note = (f"rocprofv3 FETCH_SIZE + WRITE_SIZE of k_stage2, calibrated on a "
                            f"streaming copy: profiles/{os.path.basename(TRAFFIC_REPORT)}")

    result = {
        "metric": "zone-updates/s at 512^3 HLLC+cooling (exact cooling path)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps, f64 energy recovery/accumulation",
        "data": "synthetic (seeded k^-4 lognormal clumpy medium)",
        "config": {"workload": f"clumpy self-gravitating medium {n}^3, HLLC+van Leer, exact "
                               "Z1.0 cooling, 1e4 Msun softened point mass, periodic",
                   "mesh": [n, n, n], "precision": "fp32+fp64-guard", "cooling": "exact"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_note,
                     "kernel": "stage2", "bytes_per_zone": BYTES_STAGE2,
                     "algorithmic_bytes_per_launch": BYTES_STAGE2 * zones,
                     "avg_launch_ms": s2_ms},
        "copy_probe_gbs": copy_gbs,
        "step_roofline_frac": value * BYTES_STEP / 1e9 / HBM_PEAK_GBS,
        "kernels_ms": {"stage1": s1_ms, "stage2": s2_ms},
        "nonfinite_cells": bad,
    }
    if not args.no_modes:
        result["cooling_paths"] = cooling_modes(p, u0, table, args.warmup, args.steps)
    if not args.no_gravity:
        result["self_gravity"] = self_gravity_leg(p, u0, table, args.warmup, args.steps)
    del u0
    if not args.no_l1:
        t0 = time.perf_counter()
        result["l1_vs_fp64"] = l1_fp32_vs_fp64(128, 10)
        log(f"bench: L1 check done in {time.perf_counter() - t0:.1f}s: {result['l1_vs_fp64']}")
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        zps, el = cpu_baseline(256, 20, threads)
        zps1, el1 = cpu_baseline(128, 4, 1)
        log(f"bench: CPU baseline done in {time.perf_counter() - t0:.1f}s")
        result["cpu_baseline"] = {"value": zps, "unit": "zone-updates/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"FP64 C oracle (OpenMP, {threads} threads), same "
                                            f"clumpy workload at 256^3, 20 steps ({el:.1f} s)",
                                  "host": host_cpu(),
                                  "single_thread": {"value": zps1, "cores": 1,
                                                    "sample": f"same oracle, 1 thread, clumpy 128^3, "
                                                              f"4 steps ({el1:.1f} s)"}}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
