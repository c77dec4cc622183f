"""Restatement of the learned preconditioner's network (the device's pl3d_net_*), numpy.

TEST INFRASTRUCTURE ONLY: the checker of the device's convolution, pooling, loss and gradient
kernels; nothing in the product imports it.  The reference has no learned preconditioner.

Network (pressure grid level 1; level l+1 = level l pooled 2x2x2, dims ceil(n/2)):
    x0 = [r, FLUID, SOLID, 0]
    e1 = relu(conv_a(x0))          e2 = relu(conv_b(pool(e1)))     e3 = relu(conv_c(pool(e2)))
    h2 = relu(conv_d(up(e3)))      d2 = h2 + e2
    h1 = relu(conv_e(up(d2)))      d1 = h1 + e1
    z  = conv_f(d1) * FLUID
conv: 3x3x3, zero outside the grid, no bias, W[co][tap * cin + ci] with tap = (dx+1) + 3 (dy+1)
+ 9 (dz+1) reading in[i + dx, j + dy, k + dz]; pool = (sum of the in-grid children) / 8;
up = nearest.  loss = sum_FLUID (z - t)^2 * invT2.  With bf16=True the activations and
gradient tensors are rounded to BF16 (nearest even) where the device stores them; products
and sums are exact (float64), which the device's FP32 accumulation matches to ~1e-6.
The backward pass is written out by hand and pinned by finite differences (tests).
"""
import numpy as np

LAYERS = [  # cin (padded to 4), cout, level, real input channels
    (4, 8, 0, 3), (8, 16, 1, 8), (16, 32, 2, 16), (32, 16, 1, 32), (16, 8, 0, 16), (8, 1, 0, 8)]
FLUID, SOLID = 1, 2


def kp(cin):
    return (27 * cin + 15) // 16 * 16


def unpack(params):
    """Flat device parameter array -> list of W[cout][27 cin] (padding dropped)."""
    Ws, off = [], 0
    for cin, cout, _, _ in LAYERS:
        coutp = (cout + 15) // 16 * 16
        blk = np.asarray(params[off:off + coutp * kp(cin)], np.float64).reshape(coutp, kp(cin))
        Ws.append(blk[:cout, :27 * cin].copy())
        off += coutp * kp(cin)
    return Ws


def pack(Ws):
    out = []
    for (cin, cout, _, _), W in zip(LAYERS, Ws):
        coutp = (cout + 15) // 16 * 16
        blk = np.zeros((coutp, kp(cin)))
        blk[:cout, :27 * cin] = W
        out.append(blk.ravel())
    return np.concatenate(out).astype(np.float32)


def bf16(a):
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def dims_of(n):
    d = [tuple(n)]
    for _ in range(2):
        d.append(tuple((x + 1) // 2 for x in d[-1]))
    return d


def im2col(x, dims):
    """x [nz, ny, nx, C] -> cols [V, 27 C] with column tap * C + c."""
    nx, ny, nz = dims
    C = x.shape[-1]
    xp = np.zeros((nz + 2, ny + 2, nx + 2, C))
    xp[1:-1, 1:-1, 1:-1] = x
    cols = np.empty((nz, ny, nx, 27, C))
    for tap in range(27):
        dx, dy, dz = tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1
        cols[:, :, :, tap] = xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    return cols.reshape(nz * ny * nx, 27 * C)


def conv(x, W, dims):
    nx, ny, nz = dims
    return (im2col(x, dims) @ W.T).reshape(nz, ny, nx, W.shape[0])


def conv_input_grad(g, W, dims, cin):
    """d/dx of sum(g * conv(x, W)): the transposed convolution."""
    nx, ny, nz = dims
    gc = g.reshape(-1, W.shape[0]) @ W  # [V, 27 cin]: gradient w.r.t. each column entry
    gc = gc.reshape(nz, ny, nx, 27, cin)
    out = np.zeros((nz + 2, ny + 2, nx + 2, cin))
    for tap in range(27):
        dx, dy, dz = tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1
        out[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx] += gc[:, :, :, tap]
    return out[1:-1, 1:-1, 1:-1]


def conv_weight_grad(x, g, dims):
    return g.reshape(-1, g.shape[-1]).T @ im2col(x, dims)


def pool(x, fine):
    nx, ny, nz = fine
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    xp = np.zeros((2 * cz, 2 * cy, 2 * cx, x.shape[-1]))
    xp[:nz, :ny, :nx] = x
    return xp.reshape(cz, 2, cy, 2, cx, 2, -1).sum(axis=(1, 3, 5)) / 8.0


def up(x, fine):
    nx, ny, nz = fine
    return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)[:nz, :ny, :nx]


def up_T(g, fine):
    """adjoint of up: sum over the children (= 8 pool)."""
    return pool(g, fine) * 8.0


def pool_T(g, fine):
    return up(g, fine) / 8.0


def forward(cell, r, Ws, shape, bf=True):
    """cell, r: grid arrays (nz, ny, nx) of the level-1 grid (shape = (nx, ny, nz)).
    Returns z (grid) and the cache of the backward pass."""
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    d = dims_of(shape)
    fl = (cell == FLUID).astype(np.float64)
    so = (cell == SOLID).astype(np.float64)
    x0 = np.stack([R(r * fl), fl, so, np.zeros_like(fl)], -1)
    c = {"x0": x0, "fl": fl, "d": d}
    c["e1"] = R(np.maximum(conv(x0, Ws[0], d[0]), 0))
    c["p1"] = R(pool(c["e1"], d[0]))
    c["e2"] = R(np.maximum(conv(c["p1"], Ws[1], d[1]), 0))
    c["p2"] = R(pool(c["e2"], d[1]))
    c["e3"] = R(np.maximum(conv(c["p2"], Ws[2], d[2]), 0))
    c["u3"] = up(c["e3"], d[1])
    c["h2"] = R(np.maximum(conv(c["u3"], Ws[3], d[1]), 0))
    c["d2"] = R(c["h2"] + c["e2"])
    c["u2"] = up(c["d2"], d[0])
    c["h1"] = R(np.maximum(conv(c["u2"], Ws[4], d[0]), 0))
    c["d1"] = R(c["h1"] + c["e1"])
    z = conv(c["d1"], Ws[5], d[0])[..., 0] * fl
    return z, c


def loss_and_grad(cell, r, t, invT2, Ws, shape, bf=True):
    R = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    z, c = forward(cell, r, Ws, shape, bf)
    d, fl = c["d"], c["fl"]
    diff = (z - t) * fl
    loss = float(np.sum(diff * diff) * invT2)
    gz = R(2.0 * diff * invT2)[..., None]
    g = [None] * 6
    g[5] = conv_weight_grad(c["d1"], gz, d[0])
    gd1 = conv_input_grad(gz, Ws[5], d[0], 8)
    gpe = R(gd1 * (c["h1"] > 0))
    gs1 = R(gd1)
    g[4] = conv_weight_grad(c["u2"], gpe, d[0])
    gu2 = R(conv_input_grad(gpe, Ws[4], d[0], 16))
    gd2 = up_T(gu2, d[0])
    gpd = R(gd2 * (c["h2"] > 0))
    gs2 = R(gd2)
    g[3] = conv_weight_grad(c["u3"], gpd, d[1])
    gu3 = R(conv_input_grad(gpd, Ws[3], d[1], 32))
    gpc = R(up_T(gu3, d[1]) * (c["e3"] > 0))
    g[2] = conv_weight_grad(c["p2"], gpc, d[2])
    gp2 = R(conv_input_grad(gpc, Ws[2], d[2], 16))
    gpb = R((gs2 + pool_T(gp2, d[1])) * (c["e2"] > 0))
    g[1] = conv_weight_grad(c["p1"], gpb, d[1])
    gp1 = R(conv_input_grad(gpb, Ws[1], d[1], 8))
    gpa = R((gs1 + pool_T(gp1, d[0])) * (c["e1"] > 0))
    g[0] = conv_weight_grad(c["x0"], gpa, d[0])
    return loss, z, g
