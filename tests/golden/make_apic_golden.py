"""Writes tests/golden/apic_pours.json: full gentle pours with the APIC transfer computed by the FP64
CPU oracle (the 3D reference has no APIC, so the oracle's own APIC restatement, pinned by the
affine-preservation test, is the reference for the device's APIC pours).

Usage: python tests/golden/make_apic_golden.py [cellMm ...]   (default 10 5 4; 4 mm takes ~30 min)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle3d as O  # noqa: E402
from oracle.run3d_oracle import run  # noqa: E402


def main():
    cells = [float(c) for c in sys.argv[1:]] or [10.0, 5.0, 4.0]
    path = os.path.join(HERE, "apic_pours.json")
    out = json.load(open(path)) if os.path.exists(path) else {"source": __doc__.split("\n")[0],
                                                                 "rows": []}
    O.build()
    for cell in cells:
        r = run(cell, 20.0, progress=False, transfer="apic")
        row = {"cellMm": cell, "transfer": "apic", "steps": r["result"]["steps"],
               "simulatedSeconds": r["result"]["time"],
               "dischargedMl": r["result"]["dischargedMl"],
               "spilledMl": r["result"]["spilledMl"], "lostMl": r["result"]["lostMl"],
               "dischargedFractions": r["dischargedFractions"],
               "dischargedByOrigin": r["dischargedByOrigin"],
               "pressureFailures": r["result"]["pressureFailures"],
               "wallSeconds": r["wallSeconds"]}
        out["rows"] = [x for x in out["rows"] if x["cellMm"] != cell] + [row]
        out["rows"].sort(key=lambda x: -x["cellMm"])
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
