"""The recording writer's host side (teapotfluidsimulation_amd/record3d.py): the format's own test
cases (tests/view3d.test.ts:41-67 of the reference), JSON.stringify's number rules, and the
reference's recorded header reproduced byte for byte.  No GPU."""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from teapotfluidsimulation_amd import record3d as R

from conftest import GOLDEN


def header(frames, parcels):
    """The fixture header of view3d.test.ts:11-29."""
    return {
        "version": 1, "frames": frames, "parcels": parcels, "fps": 10, "quantum": 0.00025,
        "cellMm": 5, "simulatedParcels": parcels * 3,
        "teapot": {"radius": 0.065, "height": 0.15, "wall": 0.008, "inletHeight": 0.035,
                   "boreRadius": 0.01, "spoutAngleDeg": 45, "spoutLength": 0.14},
        "pivot": [0, 0.06], "pivotWorld": [0.2, 0.2], "zOffset": 0.005,
        "domain": [0.4, 0.4, 0.08], "angles": [k * 10 for k in range(frames)],
        "poured": [[k, 2 * k, 3 * k] for k in range(frames)], "finalShares": [0.3, 0.5, 0.2],
        "label": "test",
    }


def encode_frames(h, origin, frames):
    words, prev = [], None
    for f in frames:
        w, prev = R.encode_words(f, h["quantum"], prev)
        words.append(w)
    return R.encode_recording(h, origin, words)


def test_round_trip_through_gzip_within_quantum():
    rng = np.random.default_rng(7)
    P = 50
    h = header(3, P)
    origin = rng.integers(0, 3, P).astype(np.uint8)
    fr = [rng.uniform(0, 0.3, 3 * P).astype(np.float32) for _ in range(3)]
    raw = encode_frames(h, origin, fr)
    back_h, back_o, pos = R.decode_recording(gzip.decompress(gzip.compress(raw, 9)))
    assert back_h["frames"] == 3
    assert np.array_equal(back_o, origin)
    for k in range(3):
        assert np.max(np.abs(pos[k].astype(np.float64) - fr[k])) < h["quantum"]


def test_clamps_out_of_range_positions_and_huge_jumps():
    h = header(2, 1)
    fr = [np.array([-1, 0.1, 0.1], np.float32), np.array([100, 0.1, 0.1], np.float32)]
    _, _, pos = R.decode_recording(encode_frames(h, np.zeros(1, np.uint8), fr))
    assert pos[0][0] == 0
    assert pos[1][0] == pytest.approx(32767 * h["quantum"], abs=1e-6)


def test_rejects_unknown_versions():
    h = header(1, 2)
    h["version"] = 2
    raw = encode_frames(h, np.zeros(2, np.uint8), [np.zeros(6, np.float32)])
    with pytest.raises(ValueError, match="Unsupported"):
        R.decode_recording(raw)


def test_math_round_halves_go_up():
    """Math.round: ties toward +infinity (2.5 -> 3, 3.5 -> 4), unlike numpy's half-to-even."""
    q = 0.5
    w, cur = R.encode_words(np.array([1.25, 1.75, 0.25, 2.0], np.float32), q)
    assert list(cur) == [3, 4, 1, 4]
    # running state after a clamped jump follows what was written (format.ts:71-73)
    w2, cur2 = R.encode_words(np.array([60000.0, 0, 0, 0], np.float32), 1.0, np.array([0, 5, 0, 0]))
    assert list(w2.view(np.int16)) == [32767, -5, 0, 0] and list(cur2) == [32767, 0, 0, 0]


@pytest.mark.parametrize("v,s", [
    (0.0, "0"), (-0.0, "0"), (5.0, "5"), (45, "45"), (0.00025, "0.00025"), (1e-5, "0.00001"),
    (1e-6, "0.000001"), (1e-7, "1e-7"), (1.5e-7, "1.5e-7"), (0.1 + 0.2, "0.30000000000000004"),
    (1e21, "1e+21"), (1e20, "100000000000000000000"), (123456789012345680000.0, "123456789012345680000"),
    (-2.5, "-2.5"), (2.0 ** 53, "9007199254740992"), (1.7976931348623157e308, "1.7976931348623157e+308"),
    (5e-324, "5e-324"), (float("nan"), "null"), (True, "true"),
])
def test_js_number_formats_like_ecmascript(v, s):
    assert R.js_number(v) == s


def test_reference_header_bytes_reproduced():
    """The header of the reference's recorded 5 mm pour, re-serialised, has the exact bytes (length
    and SHA-256) that scripts/record3d.ts wrote (JSON.stringify)."""
    with open(os.path.join(GOLDEN, "pour3d_5mm_header.json")) as f:
        h = json.load(f)
    with open(os.path.join(GOLDEN, "pour3d_5mm_bytes.json")) as f:
        ref = json.load(f)
    b = R.js_json(h).encode("utf-8")
    assert len(b) == ref["headerLength"]
    assert hashlib.sha256(b).hexdigest() == ref["headerSha256"]


def test_golden_frames_round_trip_through_the_encoder():
    """Decoded golden frame 0 re-encodes to the same u16 words; later kept frames, encoded as one
    recording, decode back to the stored quantised positions exactly."""
    z = np.load(os.path.join(GOLDEN, "pour3d_5mm_frames.npz"))
    q = 0.00025
    pos = z["positions"].astype(np.float64) * q  # (F, P, 3) on the quantum grid
    frames = [p.reshape(-1).astype(np.float32) for p in pos]
    w0, _ = R.encode_words(frames[0], q)
    assert np.array_equal(w0, z["positions"][0].reshape(-1))
    h = header(len(frames), z["origin"].size)
    _, o, back = R.decode_recording(encode_frames(h, z["origin"], frames))
    assert np.array_equal(o, z["origin"])
    for k in range(len(frames)):
        assert np.array_equal(np.round(back[k].astype(np.float64) / q).astype(np.int64),
                              z["positions"][k].reshape(-1).astype(np.int64))
