"""FP64 restatement of the device's multigrid preconditioner (MG-PCG), numpy + scipy.sparse.

TEST INFRASTRUCTURE ONLY: the checker of the device's k_pcg_mg; nothing in the product imports it.

The reference has no 3D multigrid; its 2D MG-PCG (src/sim/multigrid.ts:26-160) fixes the design
rules restated here in 3D:
  - coarsening by 2 per axis, a coarse cell SOLID on the border, else AIR if any child is AIR,
    else FLUID if any child is FLUID, else SOLID                       (multigrid.ts:62-79)
  - per level the same operator as the PCG system: diag = non-solid neighbours, -1 per FLUID
    neighbour, AIR = Dirichlet 0                                        (multigrid.ts:81-99, :111-123)
  - restriction = sum over the liquid children scaled to 4 x their average (the operator scales
    with the cell size squared; multigrid.ts:141-149 uses the 2D sum = 4 x average, in 3D the
    sum of 8 children is halved), prolongation piecewise constant       (multigrid.ts:150-157)
  - equal pre/post smoothing in mirrored order so the V-cycle is symmetric (multigrid.ts:1-9)
The device build replaces damped Jacobi smoothing by red-black Gauss-Seidel, red then black
going down, black then red coming up, with 2 sweeps each way and 8 + 8 mirrored sweeps on the
coarsest of 4 levels.
"""
import numpy as np
import scipy.sparse as sp

AIR, FLUID, SOLID = 0, 1, 2
LEVELS = 4
COARSE_SWEEPS = 8


def level_operator(nx, ny, nz, cell):
    """FLUID list in grid order, the 7-point matrix on it, diag and red/black colour."""
    L = np.flatnonzero(cell == FLUID)
    m = L.size
    idx = -np.ones(cell.size, np.int64)
    idx[L] = np.arange(m)
    rows, cols = [], []
    dg = np.zeros(m)
    for o in (1, -1, nx, -nx, nx * ny, -nx * ny):
        nb = cell[L + o]
        dg += nb != SOLID
        f = nb == FLUID
        rows.append(np.arange(m)[f])
        cols.append(idx[L[f] + o])
    rows = np.concatenate(rows + [np.arange(m)])
    cols = np.concatenate(cols + [np.arange(m)])
    vals = np.concatenate([-np.ones(rows.size - m), dg])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, m))
    i, j, k = L % nx, (L // nx) % ny, L // (nx * ny)
    return L, idx, A, dg, ((i + j + k) % 2).astype(np.int64)


def coarsen(nx, ny, nz, cell):
    cx, cy, cz = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
    pad = np.full((2 * cz, 2 * cy, 2 * cx), -1, np.int64)
    pad[:nz, :ny, :nx] = cell.reshape(nz, ny, nx)
    blk = pad.reshape(cz, 2, cy, 2, cx, 2).transpose(0, 2, 4, 1, 3, 5).reshape(cz, cy, cx, 8)
    air = (blk == AIR).any(-1)
    liq = (blk == FLUID).any(-1)
    out = np.where(air, AIR, np.where(liq, FLUID, SOLID)).astype(np.uint8)
    out[0], out[-1] = SOLID, SOLID
    out[:, 0], out[:, -1] = SOLID, SOLID
    out[:, :, 0], out[:, :, -1] = SOLID, SOLID
    return cx, cy, cz, out.ravel()


class Multigrid3D:
    def __init__(self, nx, ny, nz, cell):
        self.levels = []
        dims, cl = (nx, ny, nz), np.asarray(cell, np.uint8)
        for lev in range(LEVELS):
            L, idx, A, dg, col = level_operator(*dims, cl)
            if lev > 0 and L.size == 0:
                break
            self.levels.append(dict(dims=dims, L=L, idx=idx, A=A, dg=dg, col=col))
            if lev < LEVELS - 1:
                cx, cy, cz, cl = coarsen(*dims, cl)
                dims = (cx, cy, cz)
        for f, c in zip(self.levels[:-1], self.levels[1:]):
            fx, fy, _ = f["dims"]
            cx, cy, _ = c["dims"]
            i, j, k = f["L"] % fx, (f["L"] // fx) % fy, f["L"] // (fx * fy)
            par = c["idx"][(i // 2) + cx * ((j // 2) + cy * (k // 2))]
            ok = par >= 0
            f["P"] = sp.csr_matrix((np.ones(ok.sum()), (np.flatnonzero(ok), par[ok])),
                                   shape=(f["L"].size, c["L"].size))

    def _sweep(self, lv, x, b, colour):
        m = lv["col"] == colour
        r = b - lv["A"] @ x
        x[m] += np.where(lv["dg"][m] > 0, r[m] / np.where(lv["dg"][m] > 0, lv["dg"][m], 1), 0)

    def vcycle(self, b, lev=0):
        lv = self.levels[lev]
        x = np.zeros_like(b)
        if lev == len(self.levels) - 1 and lev > 0:
            for _ in range(COARSE_SWEEPS):
                self._sweep(lv, x, b, 0)
                self._sweep(lv, x, b, 1)
            for _ in range(COARSE_SWEEPS):
                self._sweep(lv, x, b, 1)
                self._sweep(lv, x, b, 0)
            return x
        for _ in range(2):
            self._sweep(lv, x, b, 0)
            self._sweep(lv, x, b, 1)
        if lev + 1 < len(self.levels):
            r = b - lv["A"] @ x
            x += lv["P"] @ self.vcycle(0.5 * (lv["P"].T @ r), lev + 1)
        for _ in range(2):
            self._sweep(lv, x, b, 1)
            self._sweep(lv, x, b, 0)
        return x


def mg_pcg(nx, ny, nz, cell, b_grid, tol, max_iter):
    """MG-PCG with the reference's stop rule max|r| <= max(tol max|b|, 1e-14)
    (solver3d.ts:902-936); returns (q on the grid, iterations, converged)."""
    mg = Multigrid3D(nx, ny, nz, cell)
    lv = mg.levels[0]
    A, L = lv["A"], lv["L"]
    b = np.asarray(b_grid, float)[L]
    q = np.zeros(cell.size)
    bn = np.abs(b).max() if b.size else 0.0
    if bn < 1e-14:
        return q, 0, True
    target = max(tol * bn, 1e-14)
    x = np.zeros_like(b)
    r = b.copy()
    z = mg.vcycle(r)
    s = z.copy()
    sig = z @ r
    for it in range(1, max_iter + 1):
        As = A @ s
        sas = s @ As
        if not sas > 0:
            break
        a = sig / sas
        x += a * s
        r -= a * As
        if np.abs(r).max() <= target:
            q[L] = x
            return q, it, True
        z = mg.vcycle(r)
        sn = z @ r
        s = z + (sn / sig) * s
        sig = sn
    q[L] = x
    return q, max_iter, False
