"""The C ABI: the library loads, exports every function include/pourlab3d.h declares, and the
ctypes records match the C layout (checked against a probe compiled from the header with gcc).
No GPU compute is called here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from teapotfluidsimulation_amd import _lib, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pourlab3d.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pl3d_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_record_layout_matches_c():
    recs = {"pl3d_teapot_config": abi.TeapotConfig, "pl3d_motion_config": abi.MotionConfig,
            "pl3d_config": abi.Config, "pl3d_diagnostics": abi.Diagnostics,
            "pl3d_info": abi.Info, "pl3d_solve_result": abi.SolveResult,
            "pl3d_poisson_result": abi.PoissonResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in recs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, py in recs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_gentle_pour_config_matches_oracle():
    from oracle import oracle3d as O
    lib = _lib.load()
    for cell in (4, 5):
        a = abi.Config()
        _lib.check(lib.pl3d_gentle_pour(cell, 20, C.byref(a)))
        b = O.gentle_pour(cell, 20)
        assert abi.config_to_dict(a) == abi.config_to_dict(b)
        assert abi.config_to_dict(abi.config_from_dict(abi.config_to_dict(a))) == abi.config_to_dict(a)


def test_error_conventions_without_gpu():
    lib = _lib.load()
    c = abi.Config()
    assert lib.pl3d_gentle_pour(-1, 20, C.byref(c)) == abi.EINVAL
    assert lib.pl3d_last_error()
    bad = abi.Config()
    h = C.c_void_p()
    assert lib.pl3d_create(C.byref(bad), C.byref(h)) == abi.EINVAL
    assert b"Config3D" in lib.pl3d_last_error()
    assert lib.pl3d_phase_name(1) == b"tilt-in"
    assert lib.pl3d_phase_name(5) == b"done"


def test_p2g_brick_table_and_tuned_settings():
    """The brick table the library exports (host-only call) and the autotuner's config merge."""
    from teapotfluidsimulation_amd import autotune
    shapes = autotune.p2g_bricks()
    assert shapes[0] == (8, 8, 4, 1024)
    for tx, ty, tz, nt in shapes:
        assert (ty + 2) * (tz + 2) <= nt and nt % 64 == 0
    cfg = {"preconditioner": "mic0", "h": 0.004}
    out = autotune.apply(cfg, {"p2gBrick": 2, "preconditioner": "chebyshev", "chebDegree": 4,
                               "chebLower": 0.1, "trials": []})
    assert out == {"preconditioner": "chebyshev", "h": 0.004, "p2gBrick": 2, "chebDegree": 4,
                   "chebLower": 0.1}
    c = abi.config_from_dict(dict(__import__("teapotfluidsimulation_amd").gentlePour3D(4, 20),
                                  **out))
    assert (c.p2g_brick, c.preconditioner, c.cheb_degree, c.cheb_lower) == (2, 3, 4, 0.1)
