"""The 2D cross-check on the device: the 3D solver (pl3d_pcg_solve) on a one-cell slab of the
reference's 2D pressure test problems (tests/pressure.test.ts) reproduces the 2D MIC(0)-PCG's
iteration counts (oracle/pressure2d.py, pinned on the CPU by tests/test_cross2d.py) within 3
(FP32 vectors; measured 21/21, 32/29), and its solution to the tolerance; the other device
preconditioners meet the same stop rule on the slab."""
import numpy as np
import pytest

from oracle import pressure2d as P
from teapotfluidsimulation_amd import Pcg3D
from tests.test_cross2d import CASES, rhs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,rows,seed,tol,cap", CASES)
def test_slab_reproduces_2d_solver(nx, ny, rows, seed, tol, cap):
    cell = P.box(nx, ny, rows)
    b = rhs(cell, seed)
    q2, it2, ok2, _ = P.solve(nx, ny, cell, b, tol, cap)
    c3, b3 = P.slab(nx, ny, cell, b)
    for pc in ("mic0", "chebyshev", "jacobi", "mg"):
        q = np.zeros(c3.size)
        r = Pcg3D(nx, ny, 3, pc).solve(c3, b3, q, tol, cap)
        qs = q.reshape(3, ny, nx)[1].ravel()
        print(f"{nx}x{ny} {pc}: {r['iterations']} iterations (2D MIC(0): {it2})")
        assert r["converged"], pc
        assert r["trueResidual"] <= tol * np.abs(b).max()
        assert np.abs(qs - q2).max() <= 1e3 * tol * np.abs(q2).max(), pc
        if pc == "mic0":  # FP32 vectors and their FP64 restarts: within 3, as in 3D
            assert abs(r["iterations"] - it2) <= 3


@pytest.mark.parametrize("nx,ny,rows,seed,tol,cap", CASES[:2])
def test_extruded_problem_has_the_2d_solution(nx, ny, rows, seed, tol, cap):
    """The 2D problem repeated over 16 layers: a grid deep enough for the multigrid's coarsening,
    so the device's MG-PCG (and the other preconditioners) is checked against the 2D MIC(0)-PCG
    solution on every layer, to ten times the solve tolerance (FP32 vectors and FP64
    refinement; measured 0.7e-9 to 1.1e-8 of max|q|)."""
    cell = P.box(nx, ny, rows)
    b = rhs(cell, seed)
    q2, _, ok2, _ = P.solve(nx, ny, cell, b, 1e-12, 2000)
    assert ok2
    nz = 18
    c3, b3 = P.extrude(nx, ny, cell, b, nz)
    for pc in ("mic0", "mg", "chebyshev", "jacobi"):
        q = np.zeros(c3.size)
        r = Pcg3D(nx, ny, nz, pc).solve(c3, b3, q, tol, 4000)
        layers = q.reshape(nz, ny, nx)[1:-1].reshape(nz - 2, -1)
        err = np.abs(layers - q2[None, :]).max() / np.abs(q2).max()
        print(f"{nx}x{ny}x{nz} {pc}: {r['iterations']} iterations, max error {err:.2e}")
        assert r["converged"], pc
        assert err <= 10 * tol, pc
