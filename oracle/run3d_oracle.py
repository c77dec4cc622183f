"""run3d-style driver for the FP64 CPU oracle (mirrors benchmarks/run3d.ts:9-38).

TEST INFRASTRUCTURE / CPU BASELINE ONLY.
Usage: python -m oracle.run3d_oracle [cellMm] [seconds] [--max-wall S] [--transfer apic] > out.json
With --max-wall the run stops after S wall seconds (a bounded CPU-baseline sample).
"""
import argparse
import json
import sys
import time

from oracle import oracle3d as O


def run(cell_mm: float, seconds: float, max_wall: float = 0.0, chunk: float = 0.5,
        progress=True, transfer: str = "flip") -> dict:
    cfg = O.gentle_pour(cell_mm, seconds)
    cfg.transfer = {"flip": 0, "apic": 1}[transfer]
    sim = O.OracleSim(cfg)
    t0 = time.perf_counter()
    rows = []
    if progress:
        sys.stderr.write(f"3D: {sim.n} parcels (half pot), grid {sim.nx}x{sim.ny}x{sim.nz}\n")
    done_at = -1.0
    truncated = False
    while not sim.finished:
        d = sim.advance(chunk)
        rows.append(d)
        if progress:
            sys.stderr.write(
                f"t={d['time']:.1f} {d['phase']} tilt={d['angleDeg']:.1f} out={d['dischargedMl']:.0f} "
                f"spill={d['spilledMl']:.0f} lost={d['lostMl']:.0f} it={d['pressureIterations']} "
                f"rho={d['interiorDensity']:.3f} wall={time.perf_counter() - t0:.0f}s\n")
            sys.stderr.flush()
        if d["phase"] == "done" and done_at < 0:
            done_at = d["time"]
        if done_at >= 0 and d["time"] > done_at + 1.5:
            break
        if max_wall and time.perf_counter() - t0 > max_wall:
            truncated = True
            break
    d = sim.diagnostics()
    order = sim.dischargeOrder()
    st = sim.state()
    counts = [int((st["origin"][order] == k).sum()) for k in range(3)]
    return {
        "cellMm": cell_mm,
        "transfer": transfer,
        "result": d,
        "dischargedFractions": sim.dischargedFractions(),
        "dischargedByOrigin": counts,
        "wallSeconds": time.perf_counter() - t0,
        "truncated": truncated,
        "history": rows,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cell", nargs="?", type=float, default=4.0)
    ap.add_argument("seconds", nargs="?", type=float, default=20.0)
    ap.add_argument("--max-wall", type=float, default=0.0)
    ap.add_argument("--transfer", choices=["flip", "apic"], default="flip")
    a = ap.parse_args()
    print(json.dumps(run(a.cell, a.seconds, a.max_wall, transfer=a.transfer), indent=1))


if __name__ == "__main__":
    main()
