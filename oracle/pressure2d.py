"""FP64 restatement of the reference's 2D MIC(0)-PCG pressure solver (src/sim/pressure.ts) and its
multigrid preconditioner (src/sim/multigrid.ts), numpy and plain loops (2D test sizes only).

TEST INFRASTRUCTURE ONLY: the checker of the 2D cross-check, in which the device's 3D solver on a
one-cell-thick slab (SOLID above and below) must reproduce this solver's iteration counts and
solutions on the reference's 2D test problems (tests/pressure.test.ts).

Followed: coefficients (pressure.ts:87-99: diag = non-solid neighbours, -1 couplings to FLUID
+x/+y neighbours), the MIC(0) factor (:178-197, tau 0.97, safety 0.25), the triangular solves in
list order (:205-222), and the CG loop with the max-norm stop rule tested on the recursive
residual after each update (:100-169).  Not followed: the gauge fixing of enclosed liquid
components (:227-268); the problems used here all touch air.  With precond="multigrid" the
preconditioner is one V-cycle of multigrid.ts: levels halved (ceil) until a side is <= 8
(:32-50); a coarse cell SOLID on the border, else AIR if any child is AIR, else FLUID if any child
is FLUID, else SOLID (:56-79); per level the diagonal counts non-solid neighbours (:80-99); two
damped-Jacobi sweeps (omega 0.8) down and up, 40 on the coarsest level, restriction = sum of the
liquid children, piecewise-constant prolongation (:104-159).
"""
import numpy as np

AIR, FLUID, SOLID = 0, 1, 2


def box(nx, ny, fluid_rows, air_top=True):
    """tests/pressure.test.ts:8-18, row-major [j][i]."""
    cell = np.zeros(nx * ny, np.uint8)
    for j in range(ny):
        for i in range(nx):
            c = i + j * nx
            if i == 0 or j == 0 or i == nx - 1 or j == ny - 1:
                cell[c] = SOLID
            elif j <= fluid_rows:
                cell[c] = FLUID
            else:
                cell[c] = AIR if air_top else SOLID
    return cell


def mg_setup(nx, ny, cell):
    """The level hierarchy (multigrid.ts:32-99): [(w, h, cell, diag, fluid list)]."""
    dims = []
    w, h = nx, ny
    while True:
        dims.append((w, h))
        if w <= 8 or h <= 8:
            break
        w, h = (w + 1) // 2, (h + 1) // 2
    levels = []
    fine = np.asarray(cell, np.uint8)
    for li, (w, h) in enumerate(dims):
        if li == 0:
            c = fine.copy()
        else:
            pw, ph, pc = levels[-1][0], levels[-1][1], levels[-1][2]
            c = np.zeros(w * h, np.uint8)
            for j in range(h):
                for i in range(w):
                    air = liquid = False
                    for dj in (0, 1):
                        for di in (0, 1):
                            fi, fj = 2 * i + di, 2 * j + dj
                            if fi >= pw or fj >= ph:
                                continue
                            t = pc[fi + fj * pw]
                            air = air or t == AIR
                            liquid = liquid or t == FLUID
                    border = i == 0 or j == 0 or i == w - 1 or j == h - 1
                    c[i + j * w] = SOLID if border else (AIR if air else (FLUID if liquid else SOLID))
        diag = np.zeros(w * h)
        lst = [k for k in range(w * h) if c[k] == FLUID]
        for k in lst:
            diag[k] = sum(c[k + o] != SOLID for o in (-1, 1, -w, w))
        levels.append((w, h, c, diag, lst))
    return levels


def mg_apply(levels, r, omega=0.8, smooth=2, coarse_sweeps=40):
    """z = M^-1 r, one V-cycle from zero (multigrid.ts:102-159)."""
    xs = [np.zeros(L[0] * L[1]) for L in levels]
    bs = [np.zeros(L[0] * L[1]) for L in levels]
    w0, h0, c0, d0, l0 = levels[0]
    for k in l0:
        bs[0][k] = r[k]

    def residual(li):
        w, h, c, diag, lst = levels[li]
        x, bb = xs[li], bs[li]
        res = np.zeros(w * h)
        for k in lst:
            ax = diag[k] * x[k]
            for o in (-1, 1, -w, w):
                if c[k + o] == FLUID:
                    ax -= x[k + o]
            res[k] = bb[k] - ax
        return res

    def jacobi(li, sweeps):
        w, h, c, diag, lst = levels[li]
        for _ in range(sweeps):
            res = residual(li)
            for k in lst:
                if diag[k] > 0:
                    xs[li][k] += omega * res[k] / diag[k]

    def vcycle(li):
        w, h, c, diag, lst = levels[li]
        for k in lst:
            xs[li][k] = 0.0
        if li == len(levels) - 1:
            jacobi(li, coarse_sweeps)
            return
        jacobi(li, smooth)
        res = residual(li)
        cw, ch, cc, cd, cl = levels[li + 1]
        for k in cl:
            bs[li + 1][k] = 0.0
        for k in lst:
            i, j = k % w, k // w
            p = (i >> 1) + (j >> 1) * cw
            if cc[p] == FLUID:
                bs[li + 1][p] += res[k]
        vcycle(li + 1)
        for k in lst:
            i, j = k % w, k // w
            p = (i >> 1) + (j >> 1) * cw
            if cc[p] == FLUID:
                xs[li][k] += xs[li + 1][p]
        jacobi(li, smooth)

    vcycle(0)
    z = np.zeros(w0 * h0)
    for k in l0:
        z[k] = xs[0][k]
    return z


def solve(nx, ny, cell, b, tol, max_iter, precond="mic0"):
    """Returns (q, iterations, converged, residual)."""
    n = nx * ny
    fl = [c for c in range(n) if cell[c] == FLUID]
    q = np.zeros(n)
    if not fl:
        return q, 0, True, 0.0
    adiag = np.zeros(n)
    ai = np.zeros(n)
    aj = np.zeros(n)
    for c in fl:
        adiag[c] = sum(cell[c + o] != SOLID for o in (-1, 1, -nx, nx))
        ai[c] = -1.0 if cell[c + 1] == FLUID else 0.0
        aj[c] = -1.0 if cell[c + nx] == FLUID else 0.0
    tau, sigma = 0.97, 0.25
    pre = np.zeros(n)
    for c in fl:
        pi = pre[c - 1] if cell[c - 1] == FLUID else 0.0
        pj = pre[c - nx] if cell[c - nx] == FLUID else 0.0
        a1, a2 = ai[c - 1], aj[c - nx]
        e = adiag[c] - (a1 * pi) ** 2 - (a2 * pj) ** 2 - tau * (
            a1 * aj[c - 1] * pi * pi + a2 * ai[c - nx] * pj * pj)
        if e < sigma * adiag[c]:
            e = adiag[c]
        pre[c] = 1.0 / np.sqrt(e) if e > 0 else 0.0

    def apply_a(x):
        y = np.zeros(n)
        for c in fl:
            y[c] = (adiag[c] * x[c] + ai[c] * x[c + 1] + ai[c - 1] * x[c - 1] + aj[c] * x[c + nx]
                    + aj[c - nx] * x[c - nx])
        return y

    mg = mg_setup(nx, ny, cell) if precond == "multigrid" else None

    def precondition(r):
        if mg is not None:
            return mg_apply(mg, r)
        qv = np.zeros(n)
        z = np.zeros(n)
        for c in fl:
            t = r[c] - ai[c - 1] * pre[c - 1] * qv[c - 1] - aj[c - nx] * pre[c - nx] * qv[c - nx]
            qv[c] = t * pre[c]
        for c in reversed(fl):
            t = qv[c] - ai[c] * pre[c] * z[c + 1] - aj[c] * pre[c] * z[c + nx]
            z[c] = t * pre[c]
        return z

    r = np.zeros(n)
    for c in fl:
        r[c] = b[c]
    bnorm = max(abs(b[c]) for c in fl)
    target = max(tol * bnorm, 1e-14)
    if bnorm <= 1e-14:
        return q, 0, True, bnorm
    z = precondition(r)
    s = z.copy()
    sig = sum(z[c] * r[c] for c in fl)
    res = bnorm
    it = 0
    for it in range(1, max_iter + 1):
        As = apply_a(s)
        sas = sum(s[c] * As[c] for c in fl)
        if not sas > 0:
            break
        alpha = sig / sas
        res = 0.0
        for c in fl:
            q[c] += alpha * s[c]
            r[c] -= alpha * As[c]
            res = max(res, abs(r[c]))
        if res <= target:
            return q, it, True, res
        z = precondition(r)
        sn = sum(z[c] * r[c] for c in fl)
        beta = sn / sig
        for c in fl:
            s[c] = z[c] + beta * s[c]
        sig = sn
    return q, min(it, max_iter), res <= target, res


def slab(nx, ny, cell2d, b2d):
    """The 2D problem as a 3D grid of depth 3 (k = 1 holds the problem, k = 0 and 2 SOLID)."""
    cell = np.full((3, ny, nx), SOLID, np.uint8)
    cell[1] = np.asarray(cell2d).reshape(ny, nx)
    b = np.zeros((3, ny, nx))
    b[1] = np.asarray(b2d).reshape(ny, nx)
    return cell.ravel(), b.ravel()


def extrude(nx, ny, cell2d, b2d, nz):
    """The 2D problem repeated on the layers k = 1 .. nz - 2 of a 3D grid (k = 0 and nz - 1
    SOLID).  For a z-uniform vector the z couplings cancel (interior layers: +2 on the diagonal,
    -1 from each layer; end layers: +1 and -1), so the 3D solution is the 2D one on every layer."""
    cell = np.full((nz, ny, nx), SOLID, np.uint8)
    b = np.zeros((nz, ny, nx))
    for k in range(1, nz - 1):
        cell[k] = np.asarray(cell2d).reshape(ny, nx)
        b[k] = np.asarray(b2d).reshape(ny, nx)
    return cell.ravel(), b.ravel()
