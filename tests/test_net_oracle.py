"""The network oracle (oracle/net3d.py) pinned without the device: its hand-written backward pass
against central finite differences of its own forward (FP64, no BF16 rounding), the parameter
packing against the device layout, and the pooling adjoints.  A dropped skip gradient, a wrong
flip in the transposed convolution or a misplaced relu mask fails the first check."""
import numpy as np

from oracle import net3d as N


def problem(shape=(7, 6, 5), seed=3):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    cell = np.full((nz, ny, nx), 2, np.uint8)
    inner = rng.choice([0, 1], size=(nz - 2, ny - 2, nx - 2), p=[0.3, 0.7]).astype(np.uint8)
    cell[1:-1, 1:-1, 1:-1] = inner
    r = rng.standard_normal((nz, ny, nx)) * (cell == 1)
    t = rng.standard_normal((nz, ny, nx)) * (cell == 1)
    Ws = [rng.standard_normal((cout, 27 * cin)) * np.sqrt(2.0 / (27 * cin)) * 0.8
          for cin, cout, _, _ in N.LAYERS]
    Ws[0][:, 3::4] = 0.0  # the zero fourth input channel
    return shape, cell, r, t, Ws


def test_backward_matches_finite_differences():
    shape, cell, r, t, Ws = problem()
    invT2 = 1.0 / np.sum(t * t)
    loss, _, g = N.loss_and_grad(cell, r, t, invT2, Ws, shape, bf=False)
    rng = np.random.default_rng(7)
    eps = 1e-6
    for layer in range(6):
        for _ in range(2):
            d = rng.standard_normal(Ws[layer].shape)
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[layer] += eps * d
            Wm[layer] -= eps * d
            lp = N.loss_and_grad(cell, r, t, invT2, Wp, shape, bf=False)[0]
            lm = N.loss_and_grad(cell, r, t, invT2, Wm, shape, bf=False)[0]
            fd = (lp - lm) / (2 * eps)
            an = float(np.sum(g[layer] * d))
            assert abs(fd - an) <= 1e-5 * max(abs(fd), abs(an), 1e-3), (layer, fd, an)
    assert loss > 0


def test_pack_roundtrip_and_padding():
    rng = np.random.default_rng(1)
    Ws = [rng.standard_normal((cout, 27 * cin)) for cin, cout, _, _ in N.LAYERS]
    p = N.pack(Ws)
    counts = [((cout + 15) // 16 * 16) * N.kp(cin) for cin, cout, _, _ in N.LAYERS]
    assert p.size == sum(counts) == 43520
    back = N.unpack(p)
    for a, b in zip(Ws, back):
        assert np.allclose(a.astype(np.float32), b)


def test_pool_up_adjoints_and_bf16():
    rng = np.random.default_rng(2)
    fine = (7, 5, 3)
    x = rng.standard_normal((3, 5, 7, 4))
    y = rng.standard_normal((2, 3, 4, 4))
    # <pool x, y> = <x, pool_T y> and <up y, x> = <y, up_T x>
    assert np.isclose(np.sum(N.pool(x, fine) * y), np.sum(x * N.pool_T(y, fine)))
    assert np.isclose(np.sum(N.up(y, fine) * x), np.sum(y * N.up_T(x, fine)))
    # BF16 rounding: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 and goes to the even 1
    assert N.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert N.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


def test_learned_helpers_operator_and_synthetic_pairs():
    """learned.apply_a is the pressure operator (oracle3d.apply_laplacian, the C++ oracle's), and
    a synthetic pair (A e, e) is consistent: solving A q = A e returns e."""
    from oracle import oracle3d as O
    from teapotfluidsimulation_amd import learned
    rng = np.random.default_rng(4)
    shape, cell, _, _, _ = problem((9, 8, 7), 5)
    nx, ny, nz = shape
    x = rng.standard_normal(cell.shape) * (cell == 1)
    ax = learned.apply_a(cell, x)
    ref = O.apply_laplacian(nx, ny, nz, cell.ravel(), x.ravel()).reshape(cell.shape)
    assert np.allclose(ax[cell == 1], ref[cell == 1], rtol=1e-14, atol=1e-14)
    c, b, e = learned.synthetic(cell, rng)
    q, res = O.pcg_solve(nx, ny, nz, cell.ravel(), b.ravel(), 1e-12, 500)
    assert np.abs(q.reshape(cell.shape) - e)[cell == 1].max() <= 1e-8 * np.abs(e).max()
