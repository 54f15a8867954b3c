"""The numpy restatement of the neural radiance cache (oracle/nrc.py), pinned by mathematics before it checks
the gfx950 kernels: the SH basis is orthonormal, the grid lookup reproduces trilinear functions exactly, and
the analytic gradients equal central finite differences."""
import numpy as np

from oracle import nrc as N


def test_sh_basis_is_orthonormal():
    # Gauss-Legendre in cos(theta) x uniform phi integrates products of bands <= 3 exactly
    ct, wt = np.polynomial.legendre.leggauss(16)
    phi = (np.arange(32) + 0.5) * 2 * np.pi / 32
    C, P = np.meshgrid(ct, phi, indexing="ij")
    st = np.sqrt(1 - C * C)
    d = np.stack([st * np.cos(P), C, st * np.sin(P)], axis=-1).reshape(-1, 3)
    w = (wt[:, None] * np.full_like(P, 2 * np.pi / 32)).reshape(-1)
    Y = N.sh16(d)
    G = (Y * w[:, None]).T @ Y
    assert np.allclose(G, np.eye(16), atol=1e-12)


def test_grid_lookup_reproduces_trilinear_functions():
    grid = np.zeros((N.LEVELS, N.T, N.FEAT))
    n = int(N.resolutions()[0])  # level 0 is dense
    xs = np.arange(n + 1)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    idx = N.grid_index(X.reshape(-1), Y.reshape(-1), Z.reshape(-1), n)
    grid[0, idx, 0] = (2 * X - 3 * Y + 0.5 * Z + 1).reshape(-1)
    grid[0, idx, 1] = (X * Y * Z).reshape(-1) * 0.01
    rng = np.random.default_rng(0)
    inp = np.zeros((200, 10))
    inp[:, :3] = rng.uniform(0, 1, (200, 3))
    inp[:, 3] = 1.0
    enc = N.encode(inp, grid, np.zeros(3), np.ones(3), rounding=False)
    q = inp[:, :3] * n
    assert np.allclose(enc[:, 0], 2 * q[:, 0] - 3 * q[:, 1] + 0.5 * q[:, 2] + 1, atol=1e-9)
    assert np.allclose(enc[:, 1], 0.01 * q[:, 0] * q[:, 1] * q[:, 2], atol=1e-9)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    n = 16
    inp = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.normal(size=(n, 3)), rng.uniform(0, 1, (n, 4))], axis=1)
    tgt = rng.uniform(0, 2, (n, 3))
    mlp = rng.normal(scale=0.3, size=N.MLP_PARAMS)
    grid = np.zeros((N.LEVELS, N.T, N.FEAT))
    lo, ie = np.zeros(3), np.ones(3)
    # give the touched grid entries values so that every layer is active
    p = N.normalise(inp, lo, ie)
    touched = []
    for lv, r in enumerate(N.resolutions()):
        for x, y, z, _ in N._corners(p, int(r)):
            idx = N.grid_index(x, y, z, int(r))
            grid[lv, idx, :] = rng.normal(size=(len(idx), 2))
            touched += [(lv, int(i)) for i in idx[:1]]
    L, gm, gg = N.loss_and_grads(inp, tgt, mlp, grid, lo, ie, rounding=False)
    y0 = N.forward(inp, mlp, grid, lo, ie, rounding=False)[0]
    den = y0 * y0 + 0.01  # the stop-gradient denominator, held fixed under the perturbations
    h = 1e-6
    for k in rng.choice(N.MLP_PARAMS, 25, replace=False):
        mp, mm = mlp.copy(), mlp.copy()
        mp[k] += h
        mm[k] -= h
        fd = (N.loss_and_grads(inp, tgt, mp, grid, lo, ie, False, den)[0] -
              N.loss_and_grads(inp, tgt, mm, grid, lo, ie, False, den)[0]) / (2 * h)
        assert abs(fd - gm[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, gm[k])
    for lv, i in touched[::4]:
        for f in range(2):
            gp, gn = grid.copy(), grid.copy()
            gp[lv, i, f] += h
            gn[lv, i, f] -= h
            fd = (N.loss_and_grads(inp, tgt, mlp, gp, lo, ie, False, den)[0] -
                  N.loss_and_grads(inp, tgt, mlp, gn, lo, ie, False, den)[0]) / (2 * h)
            assert abs(fd - gg[lv, i, f]) <= 1e-6 + 1e-4 * abs(fd), (lv, i, f, fd, gg[lv, i, f])
