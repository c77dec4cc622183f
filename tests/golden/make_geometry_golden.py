"""Writes tests/golden/geometry_pours.json: full pours of the build-only geometry and motion
variants computed by the FP64 CPU oracle (the reference has neither a rim lip nor a ramp motion,
so the oracle's restatement, pinned by the geometry and ramp tests on the CPU, checks the
device's pours).  The configurations are built here from the oracle's gentle pour, and the
library's presets (pl3d_lip_pour, pl3d_low_spout_pour) must equal them.

Usage: python tests/golden/make_geometry_golden.py [cellMm ...] [--case lip|lowSpout|ramp]
(default 10 5, every case; ~2 min per pour at 5 mm)
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle3d as O  # noqa: E402


def lip_cfg(cell_mm):
    """pl3d_lip_pour: the gentle pour from a pot with a rim lip (notch floor 12 cm, half-width
    2.5 cm, channel rising at 15 deg for 4 cm)."""
    c = O.gentle_pour(cell_mm, 20)
    c.teapot.spout_kind = 1
    c.teapot.inlet_height, c.teapot.bore_radius = 0.12, 0.025
    c.teapot.spout_angle_deg, c.teapot.spout_length = 15.0, 0.04
    return c


def low_spout_cfg(cell_mm, rate=9.0):
    """pl3d_low_spout_pour: a 6.2 cm x 14 cm pot, 2 cm tube 2 cm above the floor, CFL 1, ramp to
    90 deg at `rate` deg/s."""
    c = O.gentle_pour(cell_mm, 20)
    c.teapot.radius, c.teapot.height = 0.062, 0.14
    c.teapot.inlet_height, c.teapot.bore_radius = 0.02, 0.01
    c.cfl = 1.0
    c.motion.kind = 4
    c.motion.hold_tilt_rate = rate
    c.motion.max_angle_deg = 90.0
    return c


def ramp_cfg(cell_mm, rate=15.0):
    """pl3d_ramp_pour(cell, rate, 90): the gentle pour with a constant-rate tilt to 90 deg."""
    c = O.gentle_pour(cell_mm, 20)
    c.motion.kind = 4
    c.motion.hold_tilt_rate = rate
    c.motion.max_angle_deg = 90.0
    return c


CASES = {"lip": lip_cfg, "lowSpout": low_spout_cfg, "ramp": ramp_cfg}


def run(sim, chunk=0.5):
    """Advance until finished, or 1.5 s after the controller is done (run3d.ts's stop rule)."""
    done_at = -1.0
    while not sim.finished:
        d = sim.advance(chunk)
        if d["phase"] == "done" and done_at < 0:
            done_at = d["time"]
        if done_at >= 0 and d["time"] > done_at + 1.5:
            break
    return sim.diagnostics()


def main():
    args = sys.argv[1:]
    only = None
    if "--case" in args:  # --case NAME: only that case
        k = args.index("--case")
        only = args[k + 1]
        args = args[:k] + args[k + 2:]
    cells = [float(c) for c in args] or [10.0, 5.0]
    path = os.path.join(HERE, "geometry_pours.json")
    out = json.load(open(path)) if os.path.exists(path) else {"source": __doc__.split("\n")[0],
                                                                 "rows": []}
    O.build()
    for case, make in CASES.items():
        if only and case != only:
            continue
        for cell in cells:
            t0 = time.perf_counter()
            sim = O.OracleSim(make(cell))
            d = run(sim)
            row = {"case": case, "cellMm": cell, "parcels": sim.n, "grid": [sim.nx, sim.ny, sim.nz],
                   "steps": d["steps"], "simulatedSeconds": d["time"],
                   "dischargedMl": d["dischargedMl"], "spilledMl": d["spilledMl"],
                   "lostMl": d["lostMl"], "dischargedFractions": sim.dischargedFractions(),
                   "pressureFailures": d["pressureFailures"],
                   "wallSeconds": time.perf_counter() - t0}
            out["rows"] = [x for x in out["rows"]
                           if (x["case"], x["cellMm"]) != (case, cell)] + [row]
            out["rows"].sort(key=lambda x: (x["case"], -x["cellMm"]))
            with open(path, "w") as f:
                json.dump(out, f, indent=1)
            print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
