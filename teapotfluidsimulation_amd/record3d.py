"""Record a 3D pour for browser playback, as scripts/record3d.ts does with the CPU solver.

Usage: python -m teapotfluidsimulation_amd.record3d [cellMm] [stride] [out]

The file is the reference's pour3d.bin.gz (src/view3d/format.ts:1-8), little endian:
    u32 headerLength | header JSON (UTF-8) | u8 origin[P] | u16 frame0[3P] | i16 delta[F-1][3P]
gzip level 9.  The frames are quantised and delta-encoded on the GPU (Recorder3D), so each frame
moves 6 bytes per kept parcel to the host; the header is serialised the way JSON.stringify writes
numbers (integral values without a fraction, ECMAScript's exponent rules), so the header bytes
are those the reference script writes for the same values.
"""
import gzip
import json
import math
import struct
import sys
import time

import numpy as np

from .sim3d import Recorder3D, Simulation3D, gentlePour3D

MAX_U16 = 65535


# ---------------------------------------------------------------- JSON.stringify's number format
def js_number(v) -> str:
    """Number::toString (ECMA-262 6.1.6.1.20) of a double: the shortest round-trip digits, fixed
    notation for decimal exponents in [-6, 21), exponent notation ("1e-7", "1.5e+21") beyond."""
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        v = int(v)
        if abs(v) < 2 ** 53:
            return str(v)
        v = float(v)
    v = float(v)
    if math.isnan(v) or math.isinf(v):
        return "null"  # JSON.stringify writes non-finite numbers as null
    if v == 0:
        return "0"
    mant, _, ex = repr(abs(v)).partition("e")
    ip, _, fp = mant.partition(".")
    alldigits = ip + fp
    lead = len(alldigits) - len(alldigits.lstrip("0"))
    digits = alldigits.lstrip("0").rstrip("0") or "0"
    k = len(digits)
    n = len(ip) + (int(ex) if ex else 0) - lead  # value = 0.digits * 10^n
    if k <= n <= 21:
        out = digits + "0" * (n - k)
    elif 0 < n <= 21:
        out = digits[:n] + "." + digits[n:]
    elif -6 < n <= 0:
        out = "0." + "0" * (-n) + digits
    else:
        e = n - 1
        out = digits[0] + ("." + digits[1:] if k > 1 else "") + "e" + ("+" if e >= 0 else "-") + str(abs(e))
    return ("-" if v < 0 else "") + out


def js_json(obj) -> str:
    """JSON.stringify(obj) for dicts (insertion order), lists, strings, numbers, booleans, None."""
    if obj is None:
        return "null"
    if isinstance(obj, str):
        return json.dumps(obj, ensure_ascii=False)
    if isinstance(obj, dict):
        return "{" + ",".join(json.dumps(str(k), ensure_ascii=False) + ":" + js_json(v)
                              for k, v in obj.items()) + "}"
    if isinstance(obj, (list, tuple, np.ndarray)):
        return "[" + ",".join(js_json(v) for v in obj) + "]"
    return js_number(obj)


# ---------------------------------------------------------------- format.ts
def encode_words(positions, quantum, prev=None):
    """Host statement of one frame of encodeRecording (format.ts:56-74) from Float32 positions
    (3P): returns (the frame's 16-bit words, the running state).  `prev` None means frame 0."""
    f = np.asarray(positions, np.float32).astype(np.float64) / quantum
    r = np.floor(f)
    r = np.where(f - r >= 0.5, r + 1.0, r)  # Math.round: halves go up
    r = np.where(np.isnan(f), 0.0, r)
    cur = np.clip(r, 0, MAX_U16).astype(np.int64)
    if prev is None:
        return cur.astype(np.uint16), cur
    d = np.clip(cur - prev, -32768, 32767)
    return d.astype(np.int16).view(np.uint16), prev + d


def encode_recording(header: dict, origin, frames_words) -> bytes:
    """The raw (uncompressed) file from the header, the origin bytes and each frame's words."""
    head = js_json(header).encode("utf-8")
    parts = [struct.pack("<I", len(head)), head, np.asarray(origin, np.uint8).tobytes()]
    parts += [np.asarray(w, np.uint16).astype("<u2").tobytes() for w in frames_words]
    return b"".join(parts)


def decode_recording(buf: bytes):
    """decodeRecording (format.ts:78-99): (header, origin, positions[F][3P] Float32 metres)."""
    (hl,) = struct.unpack_from("<I", buf, 0)
    header = json.loads(buf[4:4 + hl].decode("utf-8"))
    if header.get("version") != 1:
        raise ValueError(f"Unsupported recording version {header.get('version')}")
    P = header["parcels"]
    o = 4 + hl
    origin = np.frombuffer(buf, np.uint8, P, o).copy()
    o += P
    F = header["frames"]
    words = np.frombuffer(buf, "<u2", F * 3 * P, o).reshape(F, 3 * P)
    cur = np.zeros(3 * P, np.int64)
    pos = np.zeros((F, 3 * P), np.float32)
    for k in range(F):
        cur = words[0].astype(np.int64) if k == 0 else cur + words[k].view(np.int16)
        pos[k] = (cur * header["quantum"]).astype(np.float32)
    return header, origin, pos


# ---------------------------------------------------------------- scripts/record3d.ts
def record_pour(cell_mm=5.0, stride=3, fps=10, quantum=0.00025, seconds=14.0, progress=False,
                sim_cfg=None):
    """Run the gentle pour and return (raw bytes, header).  Frame schedule and header follow
    scripts/record3d.ts:14-66: a frame at t = 0 and after every 1/fps seconds until the sim
    finishes or 1.5 s after the controller reaches 'done'."""
    cfg = dict(sim_cfg) if sim_cfg is not None else gentlePour3D(cell_mm, seconds)
    sim = Simulation3D(cfg)
    rec = Recorder3D(sim, stride, quantum)
    per_parcel_ml = sim.parcelVolume * 2 * 1e6
    words, angles, poured = [], [], []

    def capture():
        words.append(rec.capture())
        st = sim.stats()
        angles.append(sim.diagnostics()["angleDeg"])
        poured.append([c * per_parcel_ml for c in st["dischargedByOrigin"]])

    t0 = time.perf_counter()
    done_at = -1.0
    capture()
    while not sim.finished:
        sim.advance(1.0 / fps)
        capture()
        d = sim.diagnostics()
        if d["phase"] == "done" and done_at < 0:
            done_at = d["time"]
        if done_at >= 0 and d["time"] > done_at + 1.5:
            break
        if progress and len(words) % 10 == 0:
            sys.stderr.write(f"t={d['time']:.1f} poured={d['dischargedMl']:.0f} mL "
                             f"{time.perf_counter() - t0:.0f}s\n")
    shares = sim.dischargedFractions() or [0, 0, 0]
    header = {
        "version": 1,
        "frames": len(words),
        "parcels": rec.parcels,
        "fps": fps,
        "quantum": quantum,
        "cellMm": cell_mm,
        "simulatedParcels": sim.n,
        "teapot": cfg["teapot"],
        "pivot": cfg["pivot"],
        "pivotWorld": sim.pivotWorld,
        "zOffset": sim.h,
        "domain": [sim.nx * sim.h, sim.ny * sim.h, sim.nz * sim.h],
        "angles": angles,
        "poured": poured,
        "finalShares": shares,
        "label": (f"Recorded offline 3D simulation ({js_number(cell_mm)} mm cells, {sim.n} parcels "
                  f"in the half pot); every {stride}th parcel shown."),
    }
    raw = encode_recording(header, rec.origin, words)
    rec.destroy()
    sim.destroy()
    return raw, header


def main():
    args = sys.argv[1:]
    cell = float(args[0]) if len(args) > 0 else 5.0
    stride = int(args[1]) if len(args) > 1 else 3
    out = args[2] if len(args) > 2 else "pour3d.bin.gz"
    raw, h = record_pour(cell, stride, progress=True)
    gz = gzip.compress(raw, compresslevel=9)
    with open(out, "wb") as f:
        f.write(gz)
    print(f"wrote {out}: {h['frames']} frames, {h['parcels']} parcels, {len(raw) / 1e6:.1f} MB raw, "
          f"{len(gz) / 1e6:.1f} MB gzip; poured {sum(h['poured'][-1]):.1f} mL, shares "
          f"{'/'.join(f'{s * 100:.1f}' for s in h['finalShares'])} %")


if __name__ == "__main__":
    main()
