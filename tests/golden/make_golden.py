"""Write the golden fixtures under tests/golden/ from data files the reference holds.

Reads (as data only; nothing of the reference is imported or executed):
  benchmarks/run3d.json, benchmarks/gpu3d.json  -> pins.json (published CPU/GPU results)
  public/data/pour3d.bin.gz                     -> pour3d_5mm_header.json (tilt angle and poured
                                                   mL per layer per frame, pivotWorld, domain) and
                                                   pour3d_5mm_frames.npz (origin band and quantised
                                                   positions of every 3rd parcel at t = 0, 0.1, ...)
                                                   and pour3d_5mm_bytes.json (lengths and SHA-256
                                                   of the header bytes and the raw file)
The recording layout is u32 headerLength | header JSON | u8 origin[P] | u16 frame0[3P] |
i16 delta[F-1][3P], little endian, positions quantised to header.quantum metres
(src/view3d/format.ts:5-7).

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import gzip
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FRAMES_KEPT = [0, 1, 2, 3, 5, 10, 15, 20, 25, 30, 40]


def main(ref="/root/reference"):
    run3d = json.load(open(os.path.join(ref, "benchmarks/run3d.json")))
    gpu3d = json.load(open(os.path.join(ref, "benchmarks/gpu3d.json")))
    pins = {
        "source": "benchmarks/run3d.json rows and benchmarks/gpu3d.json stepComparison/rows",
        "run3d_rows": run3d["rows"],
        "gpu3d_step_comparison": gpu3d["stepComparison"],
        "gpu3d_rows": gpu3d["rows"],
    }
    with open(os.path.join(HERE, "pins.json"), "w") as f:
        json.dump(pins, f, indent=1)

    raw = gzip.open(os.path.join(ref, "public/data/pour3d.bin.gz")).read()
    hl = struct.unpack_from("<I", raw, 0)[0]
    header = json.loads(raw[4:4 + hl].decode())
    P, F = header["parcels"], header["frames"]
    o = 4 + hl
    origin = np.frombuffer(raw, np.uint8, P, o)
    o += P
    cur = np.frombuffer(raw, "<u2", 3 * P, o).astype(np.int32)
    o += 6 * P
    kept = {0: cur.copy()}
    for k in range(1, F):
        cur = cur + np.frombuffer(raw, "<i2", 3 * P, o).astype(np.int32)
        o += 6 * P
        if k in FRAMES_KEPT:
            kept[k] = cur.copy()
    with open(os.path.join(HERE, "pour3d_5mm_header.json"), "w") as f:
        json.dump(header, f, indent=1)
    # the header's exact bytes (JSON.stringify output) and the whole file's digest, so a writer
    # can be checked byte for byte against what the reference script wrote
    with open(os.path.join(HERE, "pour3d_5mm_bytes.json"), "w") as f:
        json.dump({"headerLength": hl, "headerSha256": hashlib.sha256(raw[4:4 + hl]).hexdigest(),
                   "rawLength": len(raw), "rawSha256": hashlib.sha256(raw).hexdigest()}, f, indent=1)
    np.savez_compressed(
        os.path.join(HERE, "pour3d_5mm_frames.npz"),
        origin=np.array(origin),
        frames=np.array(sorted(kept)),
        positions=np.stack([kept[k].reshape(P, 3).astype(np.uint16) for k in sorted(kept)]),
    )
    print("wrote pins.json, pour3d_5mm_header.json, pour3d_5mm_frames.npz, pour3d_5mm_bytes.json;", len(kept), "frames")


if __name__ == "__main__":
    main(*sys.argv[1:])
