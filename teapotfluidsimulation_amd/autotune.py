"""Autotuner of a pour workload: the P2G brick shape and the pressure preconditioner (with the
Chebyshev polynomial's degree and interval), chosen by measured device time.  No counterpart in
the reference (its scatter shape and Chebyshev constants are fixed, gpu3d/shaders.ts:12-21).

Each candidate runs the workload's first `probe_s` simulated seconds from the same seeded start
with per-kernel HIP-event timing; the objective is device milliseconds per simulated second over
all substep kernels (a solver setting changes the substep count slightly, so time per simulated
second, not per substep, is compared).  The search is coordinate-wise: the brick first (with the
default solver), then the preconditioner family, then the Chebyshev degree at the default
interval, then the interval at the best degree.  Every candidate solves to the same stop rule, so
the choice changes speed only; results stay within the pour's statistical parity bar.  The
degree and interval sweeps alternate until neither moves (at most `rounds` passes), since the
best degree depends on the interval.  Last, warm-started solves (each solve starts from the
previous substep's solution) are tried, and kept, with one more degree and interval sweep, when
faster.

Usage: python -m teapotfluidsimulation_amd.autotune [cellMm] [--transfer apic|flip] [--out FILE]
"""
import argparse
import ctypes as C
import json
import sys
import time

from ._lib import load
from .sim3d import Simulation3D, gentlePour3D

DEGREES = (2, 3, 4, 5, 6, 8, 10, 12)
LOWERS = (0.002, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2)
DEFAULT_SOLVER = {"preconditioner": "chebyshev", "chebDegree": 6, "chebLower": 0.05}
TUNED_KEYS = ("p2gBrick", "preconditioner", "chebDegree", "chebLower", "warmStart")


def p2g_bricks():
    """The brick shapes the library is built with: list of (tx, ty, tz, threads)."""
    lib = load()
    n = lib.pl3d_p2g_bricks(None, 0)
    buf = (C.c_int32 * (4 * n))()
    lib.pl3d_p2g_bricks(buf, n)
    return [tuple(buf[4 * b:4 * b + 4]) for b in range(n)]


def apply(cfg: dict, tuned: dict) -> dict:
    """cfg with the tuned keys of `tuned` set."""
    out = dict(cfg)
    for k in TUNED_KEYS:
        if k in tuned:
            out[k] = tuned[k]
    return out


def probe(cfg: dict, probe_s: float, chunk: float = 0.5) -> dict:
    """Device time of the first probe_s simulated seconds: {"ms_per_sim_s", "kernels", ...}."""
    sim = Simulation3D(cfg)
    try:
        sim.set_timing(True)
        t = 0.0
        while t < probe_s - 1e-9 and not sim.finished:
            d = sim.advance(min(chunk, probe_s - t))
            t = d["time"]
        kt = sim.kernel_times()
        total = sum(ms for ms, _ in kt.values())
        return {"ms_per_sim_s": total / max(t, 1e-9), "sim_s": t, "substeps": sim.steps,
                "kernels": {k: round(ms / max(t, 1e-9), 4) for k, (ms, _) in kt.items()}}
    finally:
        sim.destroy()


def tune(cfg: dict, probe_s: float = 3.0, degrees=DEGREES, lowers=LOWERS, bricks=None,
         families=("jacobi", "chebyshev"), progress=False, rounds=2, warm=True) -> dict:
    """Returns {"p2gBrick", "preconditioner", "chebDegree", "chebLower", "objective", "trials"}."""
    trials = []
    seen = {}

    def run(label, c):
        key = json.dumps({k: c.get(k) for k in TUNED_KEYS}, sort_keys=True)
        if key in seen:
            return seen[key]
        t0 = time.perf_counter()
        r = probe(c, probe_s)
        row = {"label": label, "ms_per_sim_s": round(r["ms_per_sim_s"], 3),
               "substeps": r["substeps"], "wall_s": round(time.perf_counter() - t0, 2)}
        trials.append(row)
        if progress:
            sys.stderr.write(json.dumps(row) + "\n")
            sys.stderr.flush()
        seen[key] = r["ms_per_sim_s"]
        return r["ms_per_sim_s"]

    base = apply(cfg, DEFAULT_SOLVER)
    base["p2gBrick"] = 0
    shapes = p2g_bricks()
    cand = range(len(shapes)) if bricks is None else bricks
    best_b, best = 0, None
    for b in cand:
        v = run(f"brick {shapes[b]}", dict(base, p2gBrick=b))
        if best is None or v < best:
            best_b, best = b, v
    chosen = {"p2gBrick": best_b, **DEFAULT_SOLVER}
    best_v = best
    for fam in families:
        if fam == "chebyshev":
            continue
        v = run(fam, apply(cfg, {"p2gBrick": best_b, "preconditioner": fam}))
        if v < best_v:
            chosen, best_v = {"p2gBrick": best_b, "preconditioner": fam}, v
    if "chebyshev" in families:
        cheb_best = dict(DEFAULT_SOLVER, p2gBrick=best_b)
        cheb_v = run("chebyshev default", apply(cfg, cheb_best))  # cached from the brick stage
        for _ in range(max(rounds, 1)):
            before = dict(cheb_best)
            for d in degrees:
                c = dict(cheb_best, chebDegree=d)
                v = run(f"chebyshev degree {d} lower {c['chebLower']}", apply(cfg, c))
                if v < cheb_v:
                    cheb_best, cheb_v = c, v
            for lo in lowers:
                c = dict(cheb_best, chebLower=lo)
                v = run(f"chebyshev degree {c['chebDegree']} lower {lo}", apply(cfg, c))
                if v < cheb_v:
                    cheb_best, cheb_v = c, v
            if cheb_best == before:
                break
        if cheb_v < best_v:
            chosen, best_v = cheb_best, cheb_v
    if warm:  # warm-started solves, then the Chebyshev sweeps again from there
        c = dict(chosen, warmStart=1)
        v = run("warm start", apply(cfg, c))
        if v < best_v:
            chosen, best_v = c, v
            if chosen.get("preconditioner") == "chebyshev":
                for d in degrees:
                    c = dict(chosen, chebDegree=d)
                    v = run(f"warm start, chebyshev degree {d} lower {c['chebLower']}", apply(cfg, c))
                    if v < best_v:
                        chosen, best_v = c, v
                for lo in lowers:
                    c = dict(chosen, chebLower=lo)
                    v = run(f"warm start, chebyshev degree {c['chebDegree']} lower {lo}",
                            apply(cfg, c))
                    if v < best_v:
                        chosen, best_v = c, v
    out = {k: chosen[k] for k in TUNED_KEYS if k in chosen}
    out["brickShape"] = list(shapes[out["p2gBrick"]])
    out["objective"] = {"ms_per_sim_s": round(best_v, 3), "probe_s": probe_s}
    out["trials"] = trials
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cell", nargs="?", type=float, default=4.0)
    ap.add_argument("--transfer", default="apic", choices=["apic", "flip"])
    ap.add_argument("--probe", type=float, default=3.0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    cfg = gentlePour3D(a.cell, 20)
    cfg["transfer"] = a.transfer
    res = tune(cfg, a.probe, progress=True)
    res["workload"] = {"cellMm": a.cell, "transfer": a.transfer, "preset": "gentlePour3D"}
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
