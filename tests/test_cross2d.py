"""The 2D cross-check: the reference's 2D pressure tests (tests/pressure.test.ts) against a
restatement of its 2D MIC(0)-PCG (oracle/pressure2d.py, pressure.ts), and the 3D solver
restricted to a one-cell slab (SOLID above and below: the 7-point operator becomes the 5-point
one, and 3D MIC(0) becomes 2D MIC(0)) reproducing that solver's iteration counts.  CPU: the 2D
tests' known answers, and the 3D C++ oracle on the slab equal to the 2D solver.  GPU
(test_gpu_cross2d.py): the device's 3D solve on the slab."""
import numpy as np
import pytest

from oracle import oracle3d as O
from oracle import pressure2d as P
from oracle.poisson3d import Mulberry32


def rhs(cell, seed):
    """`b[c] = rng() - 0.5` over the FLUID cells in index order (pressure.test.ts:26-27)."""
    rng = Mulberry32(seed)
    b = np.zeros(cell.size)
    for c in range(cell.size):
        if cell[c] == P.FLUID:
            b[c] = rng() - 0.5
    return b


CASES = [(20, 16, 10, 5, 1e-8, 500), (30, 30, 25, 1, 1e-8, 500), (12, 12, 9, 3, 1e-8, 300)]


def test_2d_known_answers():
    nx, ny = 20, 16
    cell = P.box(nx, ny, 10)
    b = rhs(cell, 5)
    q, it, ok, res = P.solve(nx, ny, cell, b, 1e-8, 500)
    assert ok and it < 60  # pressure.test.ts:29-41
    fl = cell == P.FLUID
    aq = O.apply_laplacian(nx, ny, 3, *P.slab(nx, ny, cell, q))[nx * ny:2 * nx * ny]
    assert np.abs(aq - b)[fl].max() < 1e-7
    # zero rhs, and an iteration cap of 2 (pressure.test.ts:73-93)
    assert P.solve(6, 6, P.box(6, 6, 3), np.zeros(36), 1e-6, 10)[1:3] == (0, True)
    c30 = P.box(30, 30, 25)
    _, it2, ok2, _ = P.solve(30, 30, c30, rhs(c30, 1), 1e-12, 2)
    assert it2 == 2 and not ok2
    # a shrunken fluid region leaves q = 0 outside it (pressure.test.ts:44-56)
    c3 = P.box(12, 12, 3)
    b3 = np.where(c3 == P.FLUID, 0.2, 0.0)
    q3, _, ok3, _ = P.solve(12, 12, c3, b3, 1e-8, 300)
    assert ok3 and np.all(q3[c3 != P.FLUID] == 0)


@pytest.mark.parametrize("nx,ny,rows,seed,tol,cap", CASES)
def test_3d_oracle_on_a_slab_is_the_2d_solver(nx, ny, rows, seed, tol, cap):
    cell = P.box(nx, ny, rows)
    b = rhs(cell, seed)
    q2, it2, ok2, _ = P.solve(nx, ny, cell, b, tol, cap)
    c3, b3 = P.slab(nx, ny, cell, b)
    q3, r3 = O.pcg_solve(nx, ny, 3, c3, b3, tol, cap)
    assert r3.iterations == it2 and ok2
    assert np.abs(q3.reshape(3, ny, nx)[1].ravel() - q2).max() <= 1e-10 * np.abs(q2).max()


@pytest.mark.parametrize("nx,ny,rows,seed,tol,cap", CASES[:2])
def test_3d_oracle_on_an_extruded_problem_has_the_2d_solution(nx, ny, rows, seed, tol, cap):
    """The 2D problem repeated over 10 layers (SOLID above and below): the 3D oracle's solution is
    the 2D solver's on every layer, to the solve tolerance."""
    cell = P.box(nx, ny, rows)
    b = rhs(cell, seed)
    q2, _, ok2, _ = P.solve(nx, ny, cell, b, 1e-12, 2000)
    nz = 12
    c3, b3 = P.extrude(nx, ny, cell, b, nz)
    q3, r = O.pcg_solve(nx, ny, nz, c3, b3, 1e-12, 4000)
    assert ok2 and r.converged
    layers = q3.reshape(nz, ny, nx)[1:-1].reshape(nz - 2, -1)
    assert np.abs(layers - q2[None, :]).max() <= 1e-9 * np.abs(q2).max()


def pool(nx, ny, rows):
    """tests/multigrid.test.ts:11-18 (the box of pressure.test.ts)."""
    return P.box(nx, ny, rows)


def test_2d_multigrid_known_answers():
    """tests/multigrid.test.ts:21-71: the hierarchy is deeper than two levels and the V-cycle is
    a symmetric operator; MG-PCG converges at 1e-8 in at most 1.5 times MIC(0)'s iterations; and
    both preconditioners give the same solution."""
    nx, ny = 40, 36
    cell = pool(nx, ny, 20)
    levels = P.mg_setup(nx, ny, cell)
    assert len(levels) > 2
    rng = Mulberry32(4)
    a, b = np.zeros(nx * ny), np.zeros(nx * ny)
    for c in range(nx * ny):
        if cell[c] == P.FLUID:
            a[c] = rng() - 0.5
            b[c] = rng() - 0.5
    ab = float(b @ P.mg_apply(levels, a))
    ba = float(a @ P.mg_apply(levels, b))
    assert abs(ab - ba) < 1e-9 * abs(ab)
    nx = ny = 60
    cell = pool(nx, ny, 45)
    rhs60 = np.zeros(nx * ny)
    for c in range(nx * ny):
        i, j = c % nx, c // nx
        if cell[c] == P.FLUID:
            rhs60[c] = np.sin(np.pi * i / nx) * np.cos(np.pi * j / 45)
    qm, itm, okm, _ = P.solve(nx, ny, cell, rhs60, 1e-8, 500, precond="mic0")
    qg, itg, okg, _ = P.solve(nx, ny, cell, rhs60, 1e-8, 500, precond="multigrid")
    print("MIC(0)", itm, "MG", itg)
    assert okm and okg and itg <= 1.5 * itm
    assert np.abs(qg - qm).max() <= 1e-6 * np.abs(qm).max()
