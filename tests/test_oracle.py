"""The FP64 CPU oracle, pinned against what the reference itself holds.

Pins (reference files): grid sizes and parcel counts (benchmarks/run3d.json:7-12, :28-33); the
CPU state after 200 substeps at 5 mm (benchmarks/gpu3d.json:14-33); the recorded 5 mm pour
(public/data/pour3d.bin.gz: pivotWorld, domain, per-frame tilt angle and poured mL per layer,
quantised positions of every 3rd parcel); the unit known answers of tests/sim3d.test.ts; closed
forms (fill height, the eased tilt-in curve, A q = b against a dense solve).
"""
import math

import numpy as np
import pytest

from oracle import oracle3d as O
from tests.conftest import quantise


def test_grid_and_parcels(pins):
    for row in pins["run3d_rows"]:
        s = O.OracleSim(O.gentle_pour(row["cellMm"], 20))
        assert [s.nx, s.ny, s.nz] == row["grid"]
        assert s.n == row["parcelsHalfPot"]
        assert max(s.bandCounts) - min(s.bandCounts) <= 1


def test_fill_height_closed_form():
    s = O.OracleSim(O.gentle_pour(8, 1))
    assert s.fillHeight == pytest.approx(1.5e-3 / (math.pi * 0.065 ** 2), rel=1e-15)
    assert s.parcelVolume * 2 * s.n == pytest.approx(1.5e-3, rel=1e-12)


def test_recording_header_and_frame0(recording):
    header, rec = recording
    s = O.OracleSim(O.gentle_pour(5, 14))
    assert s.pivotWorld == header["pivotWorld"]  # exact doubles
    assert [s.nx * s.h, s.ny * s.h, s.nz * s.h] == pytest.approx(header["domain"], rel=1e-15)
    assert s.n == header["simulatedParcels"]
    keep = np.arange(0, s.n, 3)
    st = s.state()
    # origin bands of the recorded parcels (Float32 y0 sort, ties by index)
    assert np.array_equal(st["origin"][keep], rec["origin"])
    # seeded positions (mulberry32 jitter, rejection by inCavity / teapotSdf)
    qq = quantise(st["x"][keep], st["y"][keep], st["z"][keep], header["quantum"])
    assert np.array_equal(qq, rec["positions"][0].astype(np.int64))


def test_200_substeps_and_recording_frames(pins, recording):
    """CPU block after 200 substeps (t = 0.907 s) bit-exact, then frames 1..10 of the recording."""
    cpu = pins["gpu3d_step_comparison"]["cpu"]
    s = O.OracleSim(O.gentle_pour(5, 20))
    for _ in range(200):
        s.step(min(s.stableDt(), s.cfg.duration - s.diagnostics()["time"]))
    d = s.diagnostics()
    for key in ("time", "angleDeg", "maxDivergence", "interiorDensity", "inPotMl"):
        assert d[key] == cpu[key], key
    for key in ("steps", "phase", "pressureIterations", "pressureFailures", "grid", "particles"):
        assert d[key] == cpu[key], key

    header, rec = recording
    s = O.OracleSim(O.gentle_pour(5, 14))
    keep = np.arange(0, s.n, 3)
    frames = rec["frames"]
    for f in range(1, 11):
        s.advance(0.1)
        assert s.diagnostics()["angleDeg"] == pytest.approx(header["angles"][f], abs=1e-12)
        if f in frames:
            st = s.state()
            qq = quantise(st["x"][keep], st["y"][keep], st["z"][keep], header["quantum"])
            ref = rec["positions"][frames.index(f)].astype(np.int64)
            assert np.array_equal(qq, ref), f"frame {f}"


@pytest.mark.slow
def test_first_discharge_matches_recording(recording):
    """Frames 11-22 (t = 1.1-2.24 s): the poured mL per layer of the first discharges (advect's
    spout-plane test, solver3d.ts:769-781) match the recording exactly; the tilt angle to 1e-6 deg
    (from ~frame 15 the substep lengths differ from V8's in the last bits, libm-level)."""
    header, rec = recording
    s = O.OracleSim(O.gentle_pour(5, 14))
    for _ in range(10):  # the recording's chunks (scripts/record3d.ts:40-42)
        s.advance(0.1)
    ml = s.parcelVolume * 2e6
    for f in range(11, 23):
        s.advance(0.1)
        assert s.diagnostics()["angleDeg"] == pytest.approx(header["angles"][f], abs=1e-6)
        st = s.state()
        order = s.dischargeOrder()
        poured = [int((st["origin"][order] == k).sum()) * ml for k in range(3)]
        assert poured == pytest.approx(header["poured"][f], abs=1e-9), f"frame {f}"
    assert sum(header["poured"][22]) > 5.0  # the pour has started


def test_geometry_known_answers():
    """tests/sim3d.test.ts:28-58."""
    t = O.default_teapot()
    c = t
    assert O.teapot_sdf(t, 0, 0.05, 0) > 0.03
    assert O.teapot_sdf(t, -(c.radius + c.wall / 2), 0.08, 0) < 0
    assert O.teapot_sdf(t, 0, -c.wall / 2, 0) < 0
    assert O.teapot_sdf(t, 0, c.height + 0.02, 0) > 0
    assert O.in_cavity(t, 0, 0.05, 0)
    assert not O.in_cavity(t, 0, -0.01, 0)
    p = O.build_teapot(t)
    s = 0.0
    while s < c.spout_length - 0.005:
        x, y = p["ax"] + s * p["dx"], p["ay"] + s * p["dy"]
        assert O.teapot_sdf(t, x, y, 0) > 0
        assert O.in_cavity(t, x, y, 0)
        s += 0.01
    s = 0.08
    off = c.bore_radius + c.wall / 2
    assert O.teapot_sdf(t, p["ax"] + s * p["dx"] - p["dy"] * off,
                        p["ay"] + s * p["dy"] + p["dx"] * off, 0) < 0
    assert O.teapot_sdf(t, c.radius + c.wall / 2, 0.11, 0) < 0
    assert p["tipY"] - c.bore_radius > 1.5e-3 / (math.pi * c.radius ** 2)
    # closed form: spout axis
    assert p["tipX"] == pytest.approx(0.057 + math.cos(math.pi / 4) * 0.14, abs=1e-15)


def _column(n=8):
    cell = np.zeros(n ** 3, np.uint8)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = i + n * (j + n * k)
                border = i in (0, n - 1) or j in (0, n - 1) or k in (0, n - 1)
                cell[c] = 2 if border else (1 if j < 5 else 0)
    b = np.where(cell == 1, np.sin(np.arange(n ** 3)), 0.0)
    return cell, b


def test_pcg_column_known_answers():
    """tests/sim3d.test.ts:60-80, plus A q = b checked in FP64."""
    n = 8
    cell, b = _column(n)
    q, r = O.pcg_solve(n, n, n, cell, b, 1e-8, 200)
    assert r.converged and r.iterations < 40
    assert r.true_residual <= 1e-8 * np.abs(b).max() * 1.0001
    q0, r0 = O.pcg_solve(n, n, n, cell, np.zeros(n ** 3), 1e-8, 10)
    assert r0.iterations == 0 and r0.converged
    _, r1 = O.pcg_solve(n, n, n, cell, b, 1e-14, 1)
    assert not r1.converged and r1.iterations == 1


def _dense_laplacian(nx, ny, nz, cell):
    """The 7-point operator as a dense matrix over FLUID cells (AIR neighbours: p = 0 ghost,
    SOLID: no flux), built from its definition."""
    liquid = np.flatnonzero(cell == 1)
    index = {c: t for t, c in enumerate(liquid)}
    A = np.zeros((len(liquid), len(liquid)))
    for t, c in enumerate(liquid):
        for o in (-1, 1, -nx, nx, -nx * ny, nx * ny):
            if cell[c + o] != 2:
                A[t, t] += 1
            if cell[c + o] == 1:
                A[t, index[c + o]] -= 1
    return liquid, A


def test_pcg_against_dense_solve():
    rng = np.random.default_rng(5)
    nx, ny, nz = 9, 8, 7
    cell = np.full(nx * ny * nz, 2, np.uint8)
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    flags = np.where(rng.random((nz, ny, nx)) < 0.75, 1, 0)
    flags[:, 4:, :] = np.where(rng.random((nz, ny - 4, nx)) < 0.3, 1, 0)
    flags[3, 3, 3] = 2
    cell = np.where(inner, flags, 2).astype(np.uint8).ravel()
    b = np.where(cell == 1, rng.standard_normal(cell.size), 0.0)
    liquid, A = _dense_laplacian(nx, ny, nz, cell)
    assert np.all(np.linalg.eigvalsh(A) > 0)  # SPD (every liquid component touches air)
    exact = np.linalg.solve(A, b[liquid])
    q, r = O.pcg_solve(nx, ny, nz, cell, b, 1e-12, 400)
    assert r.converged
    np.testing.assert_allclose(q[liquid], exact, rtol=0, atol=1e-9 * np.abs(exact).max())
    assert np.all(q[cell != 1] == 0)
    # the oracle's operator agrees with the dense one
    np.testing.assert_allclose(O.apply_laplacian(nx, ny, nz, cell, q)[liquid], A @ q[liquid],
                               atol=1e-12)


def test_mic0_single_cell_by_hand():
    """One FLUID cell with 5 non-solid neighbours: MIC(0) is exact (e = diag), PCG converges in
    one iteration to q = b / 5."""
    n = 5
    cell = np.full(n ** 3, 2, np.uint8)
    inner = [i + n * (j + n * k) for k in range(1, 4) for j in range(1, 4) for i in range(1, 4)]
    cell[inner] = 0
    c = 2 + n * (2 + n * 2)
    cell[c] = 1
    cell[c - n * n] = 2  # one solid neighbour below in z
    b = np.zeros(n ** 3)
    b[c] = 3.0
    q, r = O.pcg_solve(n, n, n, cell, b, 1e-12, 10)
    assert r.iterations == 1 and r.converged
    assert q[c] == pytest.approx(3.0 / 5.0, rel=1e-15)


def test_controller_tilt_in_closed_form():
    """motion.ts:121-124 with no discharge: eased tilt to 22 deg over 2 s after 0.5 s."""
    m = O.default_motion()
    c = O.Controller(m)
    t, dt = 0.0, 0.004
    while t < 2.4:
        t += dt
        a = c.advance(t, dt, 0.0)
        if t < 0.5:
            assert a == 0.0 and c.phase == "wait"
        else:
            s = min(1.0, (t - 0.5) / 2.0)
            expect = 22 * math.pi / 180 * 0.5 * (1 - math.cos(math.pi * s))
            assert a == pytest.approx(expect, abs=1e-15)
            assert c.phase == "tilt-in"
    # hold ramp +3 deg/s, then the return triggers once the target volume is reached
    for _ in range(50):
        t += dt
        c.advance(t, dt, 0.0)
    assert c.phase == "hold"
    assert c.angleDeg == pytest.approx(22 + 3 * (t - 2.5), abs=1e-9)
    t += dt
    c.advance(t, dt, 401e-6)
    assert c.phase == "return"


def _cfg8(**motion):
    c = O.gentle_pour(8, 0.3)
    c.pressure_max_iterations = 300
    c.motion.kind = 0  # still
    c.motion.target_volume = 0
    for k, v in motion.items():
        setattr(c.motion, k, v)
    return c


def test_rest_and_tilt_behaviour():
    """tests/sim3d.test.ts:83-124 on the oracle: rest, steep-tilt pour, chunk invariance."""
    s = O.OracleSim(_cfg8())
    d = s.advance(0.3)
    assert d["inPotMl"] == pytest.approx(1500, abs=1e-6)
    assert d["dischargedMl"] + d["spilledMl"] + d["lostMl"] == 0
    assert d["pressureFailures"] == 0
    assert s.dischargedFractions() is None
    assert np.all(s.state()["z"] >= s.h)

    def tilt():
        c = _cfg8(kind=1, start_delay=0, tilt_duration=0.4, hold_angle_deg=60, max_angle_deg=60,
                  max_hold=5, target_volume=400e-6)
        c.duration = 1.6
        return O.OracleSim(c)

    a, b = tilt(), tilt()
    while not a.finished:
        a.advance(0.4)
    while not b.finished:
        b.advance(0.8)
    d = a.diagnostics()
    assert d["dischargedMl"] > 20
    assert d["lostMl"] == 0
    assert d["inPotMl"] + d["dischargedMl"] + d["spilledMl"] + d["lostMl"] == pytest.approx(1500)
    assert sum(a.dischargedFractions()) == pytest.approx(1)
    sa = a.state()
    order = a.dischargeOrder()
    assert np.all(sa["status"][order] == 1)
    assert sa["y"][order].min() < a.pivotWorld[1]
    assert np.array_equal(sa["x"], b.state()["x"])
    assert np.any(sa["status"] == 0)


@pytest.mark.parametrize("fields", [
    # (velocity gradient rows, constant part): constant, general affine, rigid rotation about z
    ([[0, 0, 0], [0, 0, 0], [0, 0, 0]], [0.2, -0.1, 0.05]),
    ([[0.3, -0.4, 0.2], [0.2, 0.5, -0.1], [0.1, 0.3, -0.6]], [0.1, -0.2, 0.05]),
    ([[0, -2, 0], [2, 0, 0], [0, 0, 0]], [0, 0, 0]),
])
def test_apic_affine_preservation(fields):
    """The 3D form of the reference's APIC verification (tests/solver.test.ts:323-381): with
    v = G x + c and C = G on every parcel, particlesToGrid, validity from the transfer weights and
    gridToParticles return v and C unchanged (reference tolerances 1e-12 and 1e-11).  The
    reference checks interior parcels only; every corner of every parcel's stencil receives that
    parcel's own weight, so the check holds for all parcels here.  A wrong face offset, sign or
    corner step in either transfer breaks it for the non-constant fields."""
    G = np.array(fields[0], float)
    c0 = np.array(fields[1], float)
    cfg = O.gentle_pour(10, 20)
    cfg.transfer = 1
    sim = O.OracleSim(cfg)
    st = sim.state()
    pos = np.stack([st["x"], st["y"], st["z"]], 1)
    vel = pos @ G.T + c0
    sim.write_state(vx=vel[:, 0], vy=vel[:, 1], vz=vel[:, 2])
    sim.write_affine(G)
    sim.transfer_roundtrip()
    out = sim.state()
    got = np.stack([out["vx"], out["vy"], out["vz"]], 1)
    assert sim.n > 1000
    assert np.abs(got - vel).max() < 1e-12
    assert np.abs(sim.affine() - G).max() < 1e-11


def test_apic_is_not_flip():
    """The transfer switch changes the step: after 30 substeps the APIC and FLIP oracles differ,
    and APIC carries non-zero affine rows while FLIP keeps them zero."""
    runs = {}
    for t in (0, 1):
        cfg = O.gentle_pour(10, 20)
        cfg.transfer = t
        sim = O.OracleSim(cfg)
        for _ in range(30):
            sim.step(sim.stableDt())
        runs[t] = (sim.state(), sim.affine())
    assert np.abs(runs[0][1]).max() == 0
    assert np.abs(runs[1][1]).max() > 0
    assert np.abs(runs[0][0]["vy"] - runs[1][0]["vy"]).max() > 1e-6


def _pour_snapshot(cell_mm=10, steps=40):
    s = O.OracleSim(O.gentle_pour(cell_mm, 20))
    for _ in range(steps):
        s.step(s.stableDt())
    return s.nx, s.ny, s.nz, s.grid_field("cell").astype(np.uint8), s.grid_field("rhs_p")


def test_mg_vcycle_is_symmetric_positive_definite():
    """CG needs a symmetric positive definite preconditioner (tests/multigrid.test.ts:21-47 checks
    the 2D V-cycle the same way): <M r1, r2> = <r1, M r2> and <M r, r> > 0 on random vectors over
    a pour's liquid cells.  A smoother run in the same colour order both ways, or a restriction
    that is not a multiple of the prolongation's transpose, breaks the first."""
    from oracle.mg3d import Multigrid3D
    nx, ny, nz, cell, _ = _pour_snapshot()
    mg = Multigrid3D(nx, ny, nz, cell)
    assert len(mg.levels) >= 3
    rng = np.random.default_rng(5)
    m = mg.levels[0]["L"].size
    for _ in range(3):
        r1, r2 = rng.standard_normal(m), rng.standard_normal(m)
        a, b = mg.vcycle(r1) @ r2, r1 @ mg.vcycle(r2)
        assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
        assert mg.vcycle(r1) @ r1 > 0


def test_mg_coarsening_rule_known_answer():
    """multigrid.ts:62-79 in 3D on a hand-made 6x6x6 box: a coarse cell with an AIR child is AIR,
    one with FLUID and SOLID children only is FLUID, the coarse border is SOLID."""
    from oracle.mg3d import coarsen
    n = 6
    cell = np.full((n, n, n), 2, np.uint8)
    cell[1:5, 1:5, 1:5] = 1
    cell[3, 3, 3] = 0  # one air cell in the fine block of coarse cell (1,1,1)
    cx, cy, cz, c = coarsen(n, n, n, cell.ravel())
    assert (cx, cy, cz) == (3, 3, 3)
    c = c.reshape(3, 3, 3)
    assert c[1, 1, 1] == 0  # its children include the AIR cell (3,3,3)
    assert (c[0] == 2).all() and (c[:, 0] == 2).all() and (c[:, :, 0] == 2).all()
    cell[3, 3, 3] = 1
    _, _, _, c2 = coarsen(n, n, n, cell.ravel())
    assert c2.reshape(3, 3, 3)[1, 1, 1] == 1


def test_mg_pcg_converges_and_beats_jacobi():
    """MG-PCG meets the stop rule in far fewer iterations than Jacobi PCG on a pour snapshot, and
    its solution satisfies A q = b to the tolerance (closed check with the oracle's operator)."""
    from oracle.mg3d import mg_pcg
    nx, ny, nz, cell, b = _pour_snapshot()
    q, it, ok = mg_pcg(nx, ny, nz, cell, b, 1e-6, 400)
    assert ok and it < 40
    aq = O.apply_laplacian(nx, ny, nz, cell, q)
    assert np.abs(b - aq)[cell == 1].max() <= 1e-6 * np.abs(b[cell == 1]).max()
    _, rref = O.pcg_solve(nx, ny, nz, cell, b, 1e-6, 400)
    assert it <= rref.iterations + 2  # no worse than the reference's MIC(0)


def test_chebyshev_preconditioner_closed_form():
    """oracle/cheb3d.py's recurrence (gpu3d/shaders.ts:588-620) equals the closed form
    z = q(D^-1 A) D^-1 r, q(x) = (1 - T_k((theta - x)/delta) / T_k(theta/delta)) / x, computed by
    eigendecomposition of D^-1/2 A D^-1/2 on a pour snapshot's liquid cells; the spectrum of
    D^-1 A lies in (0, 2] and q is positive there, so the preconditioner is SPD.  Degree 1 is
    Jacobi scaled by 1/theta.  A wrong rho update, sign or D^-1 placement fails the first check."""
    from oracle.cheb3d import cheb_apply, cheb_poly
    from oracle.mg3d import level_operator
    nx, ny, nz, cell, _ = _pour_snapshot()
    # a sub-box of the snapshot keeps the dense eigenproblem small
    c3 = cell.reshape(nz, ny, nx).copy()
    c3[:, :, nx // 2:] = 2
    c3[:, ny // 2:, :] = np.where(c3[:, ny // 2:, :] == 1, 0, c3[:, ny // 2:, :])
    cell2 = c3.ravel()
    _, _, A, dg, _ = level_operator(nx, ny, nz, cell2)
    m = dg.size
    assert 200 < m < 4000
    Dh = 1.0 / np.sqrt(dg)
    S = (Dh[:, None] * A.toarray()) * Dh[None, :]
    lam, V = np.linalg.eigh(S)
    assert lam.min() > 0 and lam.max() <= 2 + 1e-12
    rng = np.random.default_rng(11)
    for degree, lower in ((6, 0.05), (3, 0.2), (9, 0.01)):
        assert np.all(cheb_poly(lam, degree, lower) > 0)
        r = rng.standard_normal(m)
        z = cheb_apply(A, dg, r, degree, lower)
        zc = Dh * (V @ (cheb_poly(lam, degree, lower) * (V.T @ (Dh * r))))
        assert np.abs(z - zc).max() <= 1e-10 * np.abs(zc).max(), (degree, lower)
    r = rng.standard_normal(m)
    assert np.allclose(cheb_apply(A, dg, r, 1, 0.05), r / (dg * 1.025), rtol=1e-15)


def test_chebyshev_pcg_known_answers():
    """The 8^3 column answers of tests/sim3d.test.ts:60-80 for the Chebyshev-PCG, its solution
    against A q = b on a pour snapshot, and fewer iterations than the reference's MIC(0)."""
    from oracle.cheb3d import cheb_pcg
    n = 8
    cell = np.zeros(n ** 3, np.uint8)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                border = i in (0, n - 1) or j in (0, n - 1) or k in (0, n - 1)
                cell[i + n * (j + n * k)] = 2 if border else (1 if j < 5 else 0)
    b = np.where(cell == 1, np.sin(np.arange(n ** 3)), 0.0)
    q, it, ok = cheb_pcg(n, n, n, cell, b, 1e-8, 200)
    assert ok and it < 40
    assert np.abs(b - O.apply_laplacian(n, n, n, cell, q))[cell == 1].max() <= 1e-8 * np.abs(b).max()
    assert cheb_pcg(n, n, n, cell, np.zeros(n ** 3), 1e-8, 10)[1] == 0
    assert cheb_pcg(n, n, n, cell, b, 1e-14, 1)[1:] == (1, False)
    nx, ny, nz, cell, b = _pour_snapshot()
    q, it, ok = cheb_pcg(nx, ny, nz, cell, b, 1e-6, 400)
    assert ok
    aq = O.apply_laplacian(nx, ny, nz, cell, q)
    assert np.abs(b - aq)[cell == 1].max() <= 1e-6 * np.abs(b[cell == 1]).max()
    _, rref = O.pcg_solve(nx, ny, nz, cell, b, 1e-6, 400)
    assert it < rref.iterations


def test_poisson_generator_oracle():
    """oracle/poisson3d.py: its mulberry32 equals the C++ oracle's (pinned by the recording), the
    separable clamped box filter equals a direct 27-point clamped mean, the flags give a liquid
    region inside the wall and under the surface, and a direct solve satisfies A x = b."""
    from oracle import poisson3d as PS
    r = PS.Mulberry32(1000)
    assert np.array_equal(np.array([r() for _ in range(16)]), O.rng(1000, 16))
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 6, 7))
    direct = np.zeros_like(a)
    for k in range(5):
        for j in range(6):
            for i in range(7):
                acc = 0.0
                for dk in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for di in (-1, 0, 1):
                            acc += a[min(max(k + dk, 0), 4), min(max(j + dj, 0), 5),
                                     min(max(i + di, 0), 6)]
                direct[k, j, i] = acc / 27
    assert np.allclose(PS.box3(a), direct, atol=1e-14)
    n = 24
    fl = PS.flags(n, 1000)
    f3 = fl.reshape(n, n, n)
    assert (f3[0] == 2).all() and (f3[:, 0] == 2).all() and (f3[:, :, 0] == 2).all()
    m = int((fl == 1).sum())
    assert 0.05 * n ** 3 < m < 0.6 * n ** 3
    b = PS.rhs(n, 1000, fl)
    assert np.all(b[fl != 1] == 0) and np.abs(b[fl == 1]).max() > 0
    x = PS.solve(n, fl, b)
    A, L = PS.operator(n, fl)
    assert np.abs(A @ x[L] - b[L]).max() < 1e-10


def test_ramp_motion_closed_form():
    """The build-only ramp motion (PL3D_MOTION_RAMP): after startDelay the tilt grows at exactly
    the ramp rate, angle = rate (t - startDelay), up to the maximum angle, where it holds; the
    reference's pour, tilt and manual kinds are unchanged (the other tests)."""
    c = O.gentle_pour(10, 20)
    c.motion.kind = 4
    c.motion.hold_tilt_rate = 45.0
    c.motion.max_angle_deg = 60.0
    c.motion.target_volume = 10.0  # never reached: no return in this window
    c.motion.max_hold = 100.0
    s = O.OracleSim(c)
    seen_hold = False
    while s.diagnostics()["time"] < 2.2:
        s.step(s.stableDt())
        d = s.diagnostics()
        t = d["time"]
        want = 0.0 if t < 0.5 else min(60.0, 45.0 * (t - 0.5))
        assert d["angleDeg"] == pytest.approx(want, abs=1e-9), t
        if want >= 60.0:
            assert d["phase"] == "hold"
            seen_hold = True
        elif t > 0.5:
            assert d["phase"] == "tilt-in"
    assert seen_hold


def _lip_teapot():
    """The rim-lip pot of pl3d_lip_pour: notch floor 12 cm, half-width 2.5 cm, lip channel
    rising at 15 deg for 4 cm."""
    t = O.default_teapot()
    t.spout_kind = 1
    t.inlet_height, t.bore_radius, t.spout_angle_deg, t.spout_length = 0.12, 0.025, 15.0, 0.04
    return t


def test_lip_geometry_closed_forms():
    """The build-only rim lip: the notch opens the rim on the pouring side only, the plate's
    and the rails' signed distances are exact at their mid-planes, and the tube's bore is gone."""
    t = _lip_teapot()
    R, w, hw, Hn, W = t.radius, t.wall, t.wall / 2, 0.12, 0.025
    p = O.build_teapot(t)
    a = math.radians(15.0)
    assert (p["ax"], p["ay"]) == (R, Hn)
    assert p["tipX"] == pytest.approx(R + 0.04 * math.cos(a), abs=1e-16)
    assert p["tipY"] == pytest.approx(Hn + 0.04 * math.sin(a), abs=1e-16)
    # the wall inside the notch is air, beside it and on the back it is solid up to the rim
    assert O.teapot_sdf(t, R + hw, 0.125, 0) > 0
    assert O.teapot_sdf(O.default_teapot(), R + hw, 0.125, 0) < 0
    z = 0.03
    assert O.teapot_sdf(t, math.sqrt((R + hw) ** 2 - z * z), 0.125, z) < 0
    assert O.teapot_sdf(t, -(R + hw), 0.145, 0) < 0
    assert O.teapot_sdf(t, R + hw, 0.115, 0) < 0  # below the notch floor
    # plate mid-plane: exactly half a wall deep; 1 mm above the face: exactly 1 mm
    u, n = (math.cos(a), math.sin(a)), (-math.sin(a), math.cos(a))

    def at(s_, h_):
        return R + s_ * u[0] + h_ * n[0], Hn + s_ * u[1] + h_ * n[1]

    assert O.teapot_sdf(t, *at(0.02, -hw), 0.01) == pytest.approx(-hw, abs=1e-15)
    assert O.teapot_sdf(t, *at(0.02, 1e-3), 0) == pytest.approx(1e-3, abs=1e-15)
    # rail mid-plane (|z| = W + w/2) at half the rim's height above the face: half a wall deep
    assert O.teapot_sdf(t, *at(0.03, 0.015), W + hw) == pytest.approx(-hw, abs=1e-15)
    # above the rails (2 mm over the rim's height) the channel is open
    assert O.teapot_sdf(t, *at(0.03, 0.15 - Hn + 2e-3), W + hw) == pytest.approx(2e-3, abs=1e-15)
    # the tube's bore is not liquid space any more; the body still is
    s = 0.08
    tube = O.build_teapot(O.default_teapot())
    bx, by = tube["ax"] + s * tube["dx"], tube["ay"] + s * tube["dy"]
    assert O.in_cavity(O.default_teapot(), bx, by, 0) and not O.in_cavity(t, bx, by, 0)
    assert O.in_cavity(t, 0.03, 0.1, 0.02) and not O.in_cavity(t, R + hw, 0.125, 0)


def test_lip_sdf_sign_matches_explicit_solid():
    """Sign of the lip pot's SDF at 40,000 random points against a membership test written
    directly from the solid's description (wall annulus up to the rim minus the notch, floor
    disk, plate slab, two rail slabs), with no min/max algebra."""
    t = _lip_teapot()
    R, w, H, Hn, W, L = t.radius, t.wall, t.height, 0.12, 0.025, 0.04
    a = math.radians(15.0)
    rng = np.random.default_rng(5)
    pts = rng.uniform([-0.09, -0.02, -0.09], [0.13, 0.18, 0.09], (40000, 3))
    mism = solid_n = 0
    for x, y, z in pts:
        r = math.hypot(x, z)
        in_notch = x > 0 and abs(z) < W and y > Hn
        wall = R <= r <= R + w and -w <= y <= H and not in_notch
        floor = -w <= y <= 0 and r <= R + w
        px, py = x - R, y - Hn
        s, above = px * math.cos(a) + py * math.sin(a), py * math.cos(a) - px * math.sin(a)
        plate = 0 <= s <= L and -w <= above <= 0 and abs(z) <= W
        rail = 0 <= s <= L and -w <= above <= H - Hn and W <= abs(z) <= W + w
        inside = wall or floor or plate or rail
        solid_n += inside
        d = O.teapot_sdf(t, x, y, z)
        if abs(d) > 1e-12:
            mism += (d < 0) != inside
    assert mism == 0 and solid_n > 1000


def test_lip_pour_behaviour():
    """Tilted by 8 deg the lip pot holds everything; held at 35 deg, liquid leaves through the
    notch and down the lip channel and is all counted as poured (none spilled)."""

    def run(angle):
        c = _cfg8(kind=2, start_delay=0, tilt_duration=1.0, hold_angle_deg=angle,
                  max_angle_deg=angle, hold_time=1.5, hold_tilt_rate=0)
        c.teapot = _lip_teapot()
        c.duration = 2.5
        s = O.OracleSim(c)
        while s.diagnostics()["time"] < 2.5:
            s.advance(0.5)
        return s.diagnostics()

    d = run(8.0)
    assert d["dischargedMl"] + d["spilledMl"] + d["lostMl"] == 0
    d = run(35.0)
    assert d["dischargedMl"] > 30 and d["spilledMl"] == 0 and d["lostMl"] == 0
    assert d["pressureFailures"] == 0
    assert d["inPotMl"] + d["dischargedMl"] == pytest.approx(1500)
