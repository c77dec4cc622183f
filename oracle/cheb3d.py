"""FP64 restatement of the Chebyshev-preconditioned CG of the reference's GPU solver
(src/gpu3d/shaders.ts:12-15 constants, :588-620 the preconditioner, :625-768 the CG loop), numpy.

TEST INFRASTRUCTURE ONLY: the checker of the device's k_pcg_cheb; nothing in the product imports it.

Preconditioner: `degree` steps of the Chebyshev iteration for D^-1 A on [lower, 2] from z = 0
(theta = (2 + lower) / 2, delta = (2 - lower) / 2, sigma1 = theta / delta):
    d = D^-1 r / theta, z = d, res = r, rho = 1 / sigma1
    repeat degree - 1 times:  res -= A d;  rho' = 1 / (2 sigma1 - rho);
                              d = rho' rho d + (2 rho' / delta) D^-1 res;  z += d;  rho = rho'
which is z = q(D^-1 A) D^-1 r with q(x) = (1 - T_k((theta - x) / delta) / T_k(theta / delta)) / x,
k = degree (the closed form the CPU tests pin).  CG: the reference's stop rule
max|r| <= max(tol max|b|, 1e-14), tested after the preconditioner as the shader does.
"""
import numpy as np

from oracle.mg3d import level_operator


def cheb_apply(A, dg, r, degree=6, lower=0.05):
    theta, delta = (2.0 + lower) / 2.0, (2.0 - lower) / 2.0
    sigma1 = theta / delta
    rho = 1.0 / sigma1
    d = r / (dg * theta)
    z = d.copy()
    res = r.copy()
    for _ in range(1, degree):
        res = res - A @ d
        rho_n = 1.0 / (2.0 * sigma1 - rho)
        d = rho_n * rho * d + (2.0 * rho_n / delta) * res / dg
        z = z + d
        rho = rho_n
    return z


def cheb_poly(x, degree=6, lower=0.05):
    """q(x) of the closed form, for x in (0, 2]."""
    theta, delta = (2.0 + lower) / 2.0, (2.0 - lower) / 2.0
    T = np.polynomial.chebyshev.Chebyshev.basis(degree)
    return (1.0 - T((theta - x) / delta) / T(theta / delta)) / x


def cheb_pcg(nx, ny, nz, cell, b_grid, tol, max_iter, degree=6, lower=0.05):
    """Returns (q on the grid, iterations, converged)."""
    L, _, A, dg, _ = level_operator(nx, ny, nz, cell)
    b = np.asarray(b_grid, float)[L]
    q = np.zeros(cell.size)
    bn = np.abs(b).max() if b.size else 0.0
    if bn < 1e-14:
        return q, 0, True
    goal = max(tol * bn, 1e-14)
    x = np.zeros_like(b)
    r = b.copy()
    z = cheb_apply(A, dg, r, degree, lower)
    s = z.copy()
    sig = z @ r
    it = 0
    while it < max_iter:
        it += 1
        As = A @ s
        sas = s @ As
        if not sas > 0:
            break
        a = sig / sas
        x += a * s
        r -= a * As
        z = cheb_apply(A, dg, r, degree, lower)
        zr = z @ r
        if np.abs(r).max() <= goal:
            q[L] = x
            return q, it, True
        s = z + (zr / sig) * s
        sig = zr
    q[L] = x
    return q, it, False
