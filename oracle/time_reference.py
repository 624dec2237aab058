"""Time the reference itself (built from its sources by oracle/ref_build.py,
serial FP64, one core) and this oracle on the same problems, in the build
container.  TEST INFRASTRUCTURE ONLY: it reports CPU figures and is never part
of the product or of a GPU run.

Per-step time = (wall time of a run of N2 steps - wall time of a run of N1
steps) / (N2 - N1), so start-up and the two output dumps cancel.  Problems:
the planar Sod tube at 128x32x32 and the point explosion at 256^3 (HLLC, van
Leer, FP64).

    python oracle/time_reference.py [--flags -O3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import ref_build  # noqa: E402
from tlaloque_amd import problems  # noqa: E402


def step_times(p, u0, n):
    o = oracle.Oracle(p)
    o.set_cons(u0)
    ts = [0.0]
    for _ in range(n):
        o.advance()
        ts.append(o.t)
    return ts


def oracle_rate(p, u0, steps, threads):
    oracle.set_threads(threads)
    o = oracle.Oracle(p)
    o.set_cons(u0)
    o.advance()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.advance()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--flags", nargs="*", default=["-O3"])
    a = ap.parse_args()
    cases = [("C1 sod 128x32x32", *problems.sod_planar(nx=128, ny=32, nz=32), 10, 30),
             ("C2 blast 256^3", *problems.blast(n=256), 2, 4)]
    out = {}
    for name, p, u0, n1, n2 in cases:
        ts = step_times(p, u0, n2 + 1)
        secs = []
        for n in (n1, n2):
            tfin = 0.5 * (ts[n] + ts[n + 1])
            r = ref_build.build_and_run("timing", p, u0, tfin, num_out=1, fflags=a.flags,
                                        timeout=3600)
            secs.append(r["seconds"])
        ref = (secs[1] - secs[0]) / (n2 - n1)
        o1 = oracle_rate(p, u0, 3, 1)
        on = oracle_rate(p, u0, 3, os.cpu_count() or 1)
        out[name] = {"zones": p.zones, "reference_s_per_step": ref,
                     "reference_zone_updates_per_s": p.zones / ref,
                     "oracle_1thread_s_per_step": o1, "oracle_1thread_zone_updates_per_s": p.zones / o1,
                     "oracle_threads": os.cpu_count(), "oracle_nthread_s_per_step": on,
                     "reference_flags": " ".join(ref_build.FFLAGS[1:] + a.flags)}
        print(name, json.dumps(out[name]), flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
