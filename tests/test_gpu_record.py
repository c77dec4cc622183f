"""Recordings on the GPU (pl3d_recorder_*, record3d.py, node/scripts/record3d.js).

- The device encoder gives exactly the words the host statement of encodeRecording
  (format.ts:56-74) gives for the device's own Float32 positions, frame after frame, including
  the u16 clamp (a quantum small enough that positions overflow 65535).
- The 5 mm recording, written as scripts/record3d.ts writes public/data/pour3d.bin.gz, against that
  file: frame 0 and the origin bands bit-exact, the header's static fields equal, the tilt angles
  of the time-scripted first second to 0.1 %, the first frames within a quantum or two, and
  the final shares and poured volume within the reference's GPU bar (verify-gpu.ts:57-61).
- The Node script writes the same bytes as the Python driver (one library, deterministic kernels).
"""
import gzip
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from teapotfluidsimulation_amd import Recorder3D, Simulation3D, gentlePour3D
from teapotfluidsimulation_amd import record3d as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "teapotfluidsimulation_amd", "node")
NODE = shutil.which("node")


def kept_positions(sim, stride):
    st = sim.state()
    return np.stack([st["x"][::stride], st["y"][::stride], st["z"][::stride]], 1).reshape(-1) \
        .astype(np.float32)


@pytest.mark.parametrize("stride,quantum,chunk,frames", [(3, 0.00025, 0.1, 16),
                                                         (1, 2e-6, 0.5, 8)])
def test_device_words_match_host_encoding(stride, quantum, chunk, frames):
    print(f"recorder stride {stride} quantum {quantum}", flush=True)
    sim = Simulation3D(gentlePour3D(5, 14))
    rec = Recorder3D(sim, stride, quantum)
    assert rec.parcels == (sim.n + stride - 1) // stride
    assert np.array_equal(rec.origin, sim.state()["origin"][::stride])
    prev = None
    clamped = 0
    for f in range(frames):
        if f:
            sim.advance(chunk)
        w = rec.capture()
        hw, prev = R.encode_words(kept_positions(sim, stride), quantum, prev)
        assert np.array_equal(w, hw), f"frame {f}"
        if f == 0:
            clamped = int(np.sum(w == 65535))
    print(f"frame-0 words clamped at 65535: {clamped}")
    if quantum < 1e-5:
        assert clamped > 0
    with pytest.raises(Exception, match="recorders"):
        sim.destroy()
    rec.destroy()
    sim.destroy()


def test_recording_5mm_matches_reference_file(recording):
    ref_h, ref = recording
    q = ref_h["quantum"]
    raw, h = R.record_pour(5, 3)
    hh, origin, pos = R.decode_recording(raw)
    assert hh == json.loads(R.js_json(h))
    for k in ("version", "fps", "quantum", "cellMm", "parcels", "simulatedParcels", "teapot",
              "pivot", "zOffset", "label"):
        assert h[k] == ref_h[k], k
    assert h["pivotWorld"] == ref_h["pivotWorld"]
    assert h["domain"] == pytest.approx(ref_h["domain"], rel=1e-15)
    assert np.array_equal(origin, ref["origin"])
    quant = [np.floor(p.astype(np.float64) / q + 0.5).astype(np.int64) for p in pos]
    assert np.array_equal(quant[0], ref["positions"][0].reshape(-1).astype(np.int64))
    for f in range(1, 11):
        # the tilt follows the script in time; the frame's time carries the substep sequence
        # (FP32 CFL speeds), which moves the small early angles in their fourth digit
        assert h["angles"][f] == pytest.approx(ref_h["angles"][f], rel=1e-3, abs=1e-7), f
    for f in (1, 2, 3, 5, 10):
        d = np.abs(quant[f] - ref["positions"][ref["frames"].index(f)].reshape(-1).astype(np.int64))
        same = float(np.mean(d == 0))
        print(f"frame {f}: {same * 100:.2f} % words identical, max |diff| {int(d.max())} quanta")
        assert same >= 0.99 and d.max() <= 2, f  # measured: >= 99.7 %, max 1 quantum
    print(f"frames {h['frames']} vs {ref_h['frames']}; poured {sum(h['poured'][-1]):.2f} vs "
          f"{sum(ref_h['poured'][-1]):.2f} mL; shares {h['finalShares']} vs {ref_h['finalShares']}")
    assert abs(h["frames"] - ref_h["frames"]) <= 2  # measured: 113 of 113
    assert abs(sum(h["poured"][-1]) / sum(ref_h["poured"][-1]) - 1) <= 0.01
    for a, b in zip(h["finalShares"], ref_h["finalShares"]):
        assert abs(a - b) <= 0.02
    assert len(gzip.compress(raw, 9)) < len(raw) / 2


@pytest.mark.skipif(NODE is None or not os.path.exists(os.path.join(NODE_DIR, "pourlab3d.node")),
                    reason="node or binding absent")
def test_node_record3d_writes_the_python_bytes(tmp_path):
    out = tmp_path / "rec.bin.gz"
    print("node record3d.js 10 2", flush=True)
    r = subprocess.run([NODE, "scripts/record3d.js", "10", "2", str(out)], cwd=NODE_DIR,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    raw, _ = R.record_pour(10, 2)
    assert gzip.decompress(out.read_bytes()) == raw
