"""FP64 restatement of the standalone pressure-Poisson problem set (the device's
pl3d_poisson_create / pl3d_poisson_solve), numpy + scipy.

TEST INFRASTRUCTURE ONLY: the checker of the device generator and batched solver.

The reference has no such set (it is build-only); its operator is the reference's pressure system:
7-point Laplacian on the FLUID cells, AIR = Dirichlet 0 ghost, SOLID = Neumann
(src/sim3d/solver3d.ts:872-881, :939-953), and its right-hand side is minus the MAC divergence of a
face velocity field (solver3d.ts:473-493) in cell units.

Problem k of a set (seed s_k = seed0 + k), on an n^3 grid, y vertical:
  draws      mulberry32 (src/sim/rng.ts:5-14) from state s_k: u0 plane height, u1/u2 tilt, u3 bump
             count, then 3 draws per bump
  surface    h(x, z) = (0.3 + 0.3 u0) n + tx (x - n/2) + tz (z - n/2)
                      + sum_b a_b exp(-((x - cx_b)^2 + (z - cz_b)^2) / (2 * 6^2)),
             tx, tz = 0.2 (u - 0.5), 4 + floor(5 u3) bumps with centres (u n, u n), amplitude
             8 (2u - 1) cells, evaluated at cell centres
  flags      SOLID on the box border and where the cell centre's distance from the vertical axis
             is >= n/2 - 3 (a 2-cell wall whose outer radius is n/2 - 1, and everything outside);
             FLUID where the centre's height is below h; else AIR
  velocity   independent N(0, 1) on every face from a counter-based hash: u = mulberry32 step of the
             state seed*0x9E3779B1 + index*0x85EBCA77 (mod 2^32), Box-Muller with two such draws
             per face; then a 3x3x3 box filter per component (indices clamped to the component's
             grid)
  rhs        -(u[i+1] - u[i] + v[j+1] - v[j] + w[k+1] - w[k]) on FLUID cells, 0 elsewhere, stored
             as FP32
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

AIR, FLUID, SOLID = 0, 1, 2
M32 = 0xFFFFFFFF


def mulberry_step(state):
    """One mulberry32 step (rng.ts:5-14) from each uint state; returns uniforms in [0, 1)."""
    a = (np.asarray(state, np.uint64) + 0x6D2B79F5) & M32
    t = a
    t = ((t ^ (t >> 15)) * (t | 1)) & M32
    t = (t ^ ((t + ((t ^ (t >> 7)) * (t | 61))) & M32)) & M32
    return ((t ^ (t >> 14)) & M32).astype(np.float64) / 4294967296.0


class Mulberry32:
    def __init__(self, seed):
        self.a = seed & M32

    def __call__(self):
        u = float(mulberry_step(np.uint64(self.a)))
        self.a = (self.a + 0x6D2B79F5) & M32
        return u


def hash_uniform(seed, idx):
    state = (np.uint64(seed) * np.uint64(0x9E3779B1) + np.asarray(idx, np.uint64) *
             np.uint64(0x85EBCA77)) & np.uint64(M32)
    return mulberry_step(state)


def face_normals(seed, count, offset=0):
    idx = np.arange(count, dtype=np.uint64) + np.uint64(offset)
    u1 = hash_uniform(seed, 2 * idx)
    u2 = hash_uniform(seed, 2 * idx + 1)
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)


def surface(n, seed):
    rng = Mulberry32(seed)
    H = (0.3 + 0.3 * rng()) * n
    tx, tz = 0.2 * (rng() - 0.5), 0.2 * (rng() - 0.5)
    nb = 4 + int(np.floor(5 * rng()))
    bumps = [(rng() * n, rng() * n, 8.0 * (2.0 * rng() - 1.0)) for _ in range(nb)]
    return H, tx, tz, bumps


def flags(n, seed):
    H, tx, tz, bumps = surface(n, seed)
    c = np.arange(n) + 0.5
    X, Z = np.meshgrid(c, c, indexing="xy")  # [z, x]
    h = H + tx * (X - n / 2) + tz * (Z - n / 2)
    for cx, cz, a in bumps:
        h = h + a * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2.0 * 36.0))
    r = np.sqrt((X - n / 2) ** 2 + (Z - n / 2) ** 2)
    f = np.empty((n, n, n), np.uint8)  # [k(z), j(y), i(x)]
    y = c[None, :, None]
    f[:] = np.where(y < h[:, None, :], FLUID, AIR)
    f[np.broadcast_to((r >= n / 2 - 3)[:, None, :], f.shape)] = SOLID
    f[0], f[-1] = SOLID, SOLID
    f[:, 0], f[:, -1] = SOLID, SOLID
    f[:, :, 0], f[:, :, -1] = SOLID, SOLID
    return f.ravel()


def box3(a):
    """3x3x3 box filter with clamped indices, a indexed [k, j, i]."""
    out = np.zeros_like(a)
    nk, nj, ni = a.shape
    for dk in (-1, 0, 1):
        kk = np.clip(np.arange(nk) + dk, 0, nk - 1)
        for dj in (-1, 0, 1):
            jj = np.clip(np.arange(nj) + dj, 0, nj - 1)
            for di in (-1, 0, 1):
                ii = np.clip(np.arange(ni) + di, 0, ni - 1)
                out += a[kk][:, jj][:, :, ii]
    return out / 27.0


def rhs(n, seed, fl=None):
    fl = flags(n, seed) if fl is None else fl
    nu = (n + 1) * n * n
    raw = face_normals(seed, 3 * nu)
    u = box3(raw[:nu].reshape(n, n, n + 1))
    v = box3(raw[nu:2 * nu].reshape(n, n + 1, n))
    w = box3(raw[2 * nu:].reshape(n + 1, n, n))
    div = (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])
    b = np.where(fl.reshape(n, n, n) == FLUID, -div, 0.0)
    return b.astype(np.float32).ravel()


def operator(n, fl):
    """The pressure matrix on the FLUID cells (grid order) and the list of FLUID cells."""
    L = np.flatnonzero(fl == FLUID)
    m = L.size
    idx = -np.ones(fl.size, np.int64)
    idx[L] = np.arange(m)
    rows, cols = [], []
    dg = np.zeros(m)
    for o in (1, -1, n, -n, n * n, -n * n):
        nb = fl[L + o]
        dg += nb != SOLID
        f = nb == FLUID
        rows.append(np.arange(m)[f])
        cols.append(idx[L[f] + o])
    rows = np.concatenate(rows + [np.arange(m)])
    cols = np.concatenate(cols + [np.arange(m)])
    vals = np.concatenate([-np.ones(rows.size - m), dg])
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, m)), L


def solve(n, fl, b, tol=1e-12):
    """FP64 solution on the grid (direct sparse factorisation for the small test sizes)."""
    A, L = operator(n, fl)
    x = np.zeros(fl.size)
    if L.size:
        x[L] = spla.spsolve(A.tocsc(), np.asarray(b, np.float64)[L])
    return x
