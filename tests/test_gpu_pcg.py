"""Device PCG (pl3d_pcg_solve, one workgroup) against the FP64 oracle's Pcg3D.solve
(solver3d.ts:858-937).

Tolerances: the device solve stops only when its FP64 true residual max|b - A q| passes the
reference's rule max|r| <= max(tol * max|b|, 1e-14); q then differs from the oracle's q by at most
what two tol-accurate solutions can differ, checked as |q - q_ref|_inf <= 1e3 * tol * |q_ref|_inf.
MIC(0) is the same preconditioner as the oracle's (hyperplane order = lexicographic order for this
stencil), MG the same V-cycle as oracle/mg3d.py and Chebyshev the same polynomial as
oracle/cheb3d.py (the reference's GPU preconditioner), so iteration counts agree to within a few
(FP32 vectors).
"""
import numpy as np
import pytest

from oracle import oracle3d as O
from oracle.cheb3d import cheb_pcg
from oracle.mg3d import mg_pcg
from teapotfluidsimulation_amd import Pcg3D

pytestmark = pytest.mark.gpu


def column(n=8):
    cell = np.zeros(n ** 3, np.uint8)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = i + n * (j + n * k)
                border = i in (0, n - 1) or j in (0, n - 1) or k in (0, n - 1)
                cell[c] = 2 if border else (1 if j < 5 else 0)
    b = np.where(cell == 1, np.sin(np.arange(n ** 3)), 0.0)
    return cell, b


def random_domain(nx, ny, nz, seed, fluid=0.7):
    rng = np.random.default_rng(seed)
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    flags = rng.choice([0, 1, 2], size=(nz, ny, nx), p=[1 - fluid - 0.05, fluid, 0.05])
    top = ny * 2 // 3
    flags[:, top:, :] = np.where(rng.random((nz, ny - top, nx)) < 0.2, 1, 0)
    cell = np.where(inner, flags, 2).astype(np.uint8).ravel()
    b = np.where(cell == 1, rng.standard_normal(cell.size), 0.0)
    return cell, b


@pytest.mark.parametrize("precond", ["mic0", "jacobi", "mg", "chebyshev"])
def test_column_known_answers(precond):
    """tests/sim3d.test.ts:60-80 on the device."""
    n = 8
    cell, b = column(n)
    pcg = Pcg3D(n, n, n, precond)
    q = np.zeros(n ** 3)
    r = pcg.solve(cell, b, q, 1e-8, 200)
    assert r["converged"] and r["iterations"] < 40
    assert r["trueResidual"] <= 1e-8 * np.abs(b).max()
    aq = O.apply_laplacian(n, n, n, cell, q)
    assert np.abs(b - aq)[cell == 1].max() <= 1e-8 * np.abs(b).max()
    assert pcg.solve(cell, np.zeros(n ** 3), q, 1e-8, 10)["iterations"] == 0
    r1 = pcg.solve(cell, b, q, 1e-14, 1)
    assert not r1["converged"] and r1["iterations"] == 1


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("shape", [(12, 10, 9), (24, 20, 14)])
def test_random_domains_vs_oracle(seed, shape):
    nx, ny, nz = shape
    cell, b = random_domain(nx, ny, nz, seed)
    tol = 1e-6
    qref, rref = O.pcg_solve(nx, ny, nz, cell, b, tol, 400)
    for precond in ("mic0", "jacobi", "mg", "chebyshev"):
        q = np.zeros(nx * ny * nz)
        r = Pcg3D(nx, ny, nz, precond).solve(cell, b, q, tol, 400)
        assert r["converged"], precond
        aq = O.apply_laplacian(nx, ny, nz, cell, q)
        true_res = np.abs(b - aq)[cell == 1].max()
        assert true_res <= tol * np.abs(b).max(), precond
        assert r["trueResidual"] == pytest.approx(true_res, rel=1e-6, abs=1e-15)
        assert np.abs(q - qref).max() <= 1e3 * tol * np.abs(qref).max(), precond
        assert np.all(q[cell != 1] == 0)
        if precond == "mic0":
            assert abs(r["iterations"] - rref.iterations) <= 3
        if precond == "mg":  # same V-cycle as oracle/mg3d.py (FP32 vectors on the device)
            assert abs(r["iterations"] - mg_pcg(nx, ny, nz, cell, b, tol, 400)[1]) <= 3
        if precond == "chebyshev":
            assert abs(r["iterations"] - cheb_pcg(nx, ny, nz, cell, b, tol, 400)[1]) <= 3


def test_empty_and_capacity():
    nx = ny = nz = 6
    cell = np.full(nx * ny * nz, 0, np.uint8)
    q = np.zeros(nx * ny * nz)
    r = Pcg3D(nx, ny, nz).solve(cell, np.ones(nx * ny * nz), q, 1e-6, 10)
    assert r["iterations"] == 0 and r["converged"] and r["liquidCells"] == 0


@pytest.mark.parametrize("precond", ["mic0", "mg", "chebyshev"])
def test_pour_snapshot_vs_oracle(precond):
    """(cell, rhs) of a 5 mm pour substep taken from the oracle, solved on the device."""
    s = O.OracleSim(O.gentle_pour(5, 20))
    for _ in range(30):
        s.step(s.stableDt())
    cell = s.grid_field("cell").astype(np.uint8)
    b = s.grid_field("rhs_p")
    qref = s.grid_field("q")
    it_ref = s.last_solves()[0]
    q = np.zeros_like(b)
    r = Pcg3D(s.nx, s.ny, s.nz, precond).solve(cell, b, q, 1e-6, 400)
    assert r["converged"]
    assert r["trueResidual"] <= 1e-6 * np.abs(b).max()
    if precond == "mic0":
        assert abs(r["iterations"] - it_ref) <= 3
    elif precond == "mg":
        assert abs(r["iterations"] - mg_pcg(s.nx, s.ny, s.nz, cell, b, 1e-6, 400)[1]) <= 3
    else:
        assert abs(r["iterations"] - cheb_pcg(s.nx, s.ny, s.nz, cell, b, 1e-6, 400)[1]) <= 3
    assert np.abs(q - qref).max() <= 1e-3 * np.abs(qref).max()
