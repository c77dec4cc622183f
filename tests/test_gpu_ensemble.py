"""Ensembles: independent pours on different grids stepped by one launch per kernel.  Bar: every
member's diagnostics, parcel state and discharge shares are bit-identical to the same config run
alone as a Simulation3D and advanced by the same durations (the kernels are deterministic and a
member's blocks touch only that member's memory)."""
import ctypes as C

import numpy as np
import pytest

from teapotfluidsimulation_amd import Ensemble3D, Simulation3D, gentlePour3D, variedPour3D
from teapotfluidsimulation_amd._lib import load

pytestmark = pytest.mark.gpu


def configs():
    cfgs = [variedPour3D(7, i) for i in range(3)] + [gentlePour3D(10, 20)]
    cfgs[1]["transfer"] = "apic"
    for c in cfgs:
        c["preconditioner"] = "jacobi"
    return cfgs


def test_members_match_single_runs():
    cfgs = configs()
    print("ensemble of", len(cfgs), flush=True)
    ens = Ensemble3D(cfgs)
    grids = {(m.nx, m.ny, m.nz) for m in ens.members}
    assert len(grids) == len(cfgs)  # uneven grids
    ens_diag = [ens.advance(0.5) for _ in range(4)]
    ens_state = [m.state() for m in ens.members]
    ens_frac = [m.dischargedFractions() for m in ens.members]
    ens_stats = [m.stats() for m in ens.members]
    ens.destroy()
    for i, cfg in enumerate(cfgs):
        print("single", i, flush=True)
        sim = Simulation3D(cfg)
        for k in range(4):
            d = sim.advance(0.5)
            assert d == ens_diag[k][i], (i, k)
        st = sim.state()
        for key in ("x", "y", "z", "vx", "vy", "vz", "status"):
            assert np.array_equal(st[key], ens_state[i][key]), (i, key)
        assert sim.dischargedFractions() == ens_frac[i]
        assert ens_stats[i]["solverOverflow"] == 0 and ens_stats[i]["pressureFailures"] == 0
        sim.destroy()


def test_mask_holds_members():
    cfgs = configs()[:3]
    ens = Ensemble3D(cfgs)
    d1 = ens.advance(0.25, mask=[True, False, True])
    assert d1[1]["steps"] == 0 and d1[0]["steps"] > 0 and d1[2]["steps"] > 0
    d2 = ens.advance(0.25, mask=[False, True, False])
    assert d2[0]["steps"] == d1[0]["steps"] and d2[1]["time"] >= 0.25 - 1e-9
    ens.destroy()


def test_member_handles_are_not_steppable():
    ens = Ensemble3D(configs()[:2])
    lib = load()
    h = ens.members[0]._h
    assert lib.pl3d_step(h, C.c_double(1e-3)) == -4
    assert lib.pl3d_advance(h, C.c_double(0.1), None) == -4
    assert lib.pl3d_destroy(h) == -4
    d = ens.members[0].diagnostics()  # queries work
    assert d["steps"] == 0
    ens.destroy()


def test_mixed_preconditioners_rejected():
    cfgs = configs()[:2]
    cfgs[1]["preconditioner"] = "mic0"
    with pytest.raises(RuntimeError, match="preconditioner"):
        Ensemble3D(cfgs)
