"""Pins the learned-form oracle (oracle/denoiser.py) by definitions and properties, not by
restating its formulas: a brute-force single-voxel sum, shift equivariance, adjointness, finite
differences, torch's own BF16 rounding, and the discrete continuity the projection restores."""
import numpy as np
import pytest

from oracle import denoiser as D


def _rand(rng, *shape):
    return rng.standard_normal(shape)


def test_conv_matches_brute_force_sum():
    rng = np.random.default_rng(1)
    x = _rand(rng, 3, 3, 4, 5)
    W = _rand(rng, 27, 3, 2)
    b = _rand(rng, 2)
    out = D.conv3d_periodic(x, W, b)
    n = x.shape[1:]
    for (i, j, k) in [(0, 0, 0), (2, 3, 4), (1, 0, 2)]:
        for co in range(2):
            s = b[co]
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        t = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)
                        for ci in range(3):
                            s += W[t, ci, co] * x[ci, (i + dx) % n[0], (j + dy) % n[1], (k + dz) % n[2]]
            assert abs(out[co, i, j, k] - s) < 1e-12


def test_conv_is_shift_equivariant():
    rng = np.random.default_rng(2)
    x = _rand(rng, 4, 5, 6, 7)
    W = _rand(rng, 27, 4, 3)
    a = D.conv3d_periodic(np.roll(x, (2, -1, 3), axis=(1, 2, 3)), W)
    b = np.roll(D.conv3d_periodic(x, W), (2, -1, 3), axis=(1, 2, 3))
    assert np.abs(a - b).max() < 1e-12


def test_transpose_and_weight_grad_are_adjoints():
    rng = np.random.default_rng(3)
    x = _rand(rng, 5, 4, 6, 5)
    g = _rand(rng, 3, 4, 6, 5)
    W = _rand(rng, 27, 5, 3)
    lhs = np.sum(D.conv3d_periodic(x, W) * g)
    assert abs(lhs - np.sum(x * D.conv3d_transpose(g, W))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(W * D.weight_grad(x, g))) < 1e-9 * abs(lhs)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 5000),
                        np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 3.0e38])])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(D.bf16(v), ref)


@pytest.mark.parametrize("margin,n", [(0, (4, 4, 4)), (1, (5, 4, 6))])
def test_net_gradient_matches_finite_differences(margin, n):
    """central differences of the loss, over all voxels and over the interior of a crop"""
    rng = np.random.default_rng(5)
    p = D.init_params(7).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.shape)          # non-zero biases too
    x = _rand(rng, 3, *n)
    y = _rand(rng, 3, *n)
    loss, g = D.net_grad(p, x, y, emulate=False, margin=margin)
    assert abs(loss - D.net_loss(p, x, y, False, margin)) < 1e-12 * loss
    idx = rng.choice(D.NPARAMS, 40, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (D.net_loss(pp, x, y, False, margin) - D.net_loss(pm, x, y, False, margin)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_margin_loss_counts_only_the_interior():
    """margin 0 is the plain mean over every voxel; with a margin the target outside the interior
    does not enter the loss or the gradient, and the loss is the mean over the interior alone"""
    rng = np.random.default_rng(8)
    p = D.init_params(3).astype(np.float64)
    x = _rand(rng, 3, 7, 8, 9)
    y = _rand(rng, 3, 7, 8, 9)
    out, _ = D.net_forward(p, x, False)
    assert abs(D.net_loss(p, x, y, False) - np.mean((out - y) ** 2)) < 1e-12
    l1, g1 = D.net_grad(p, x, y, False, margin=2)
    y2 = y.copy()
    y2[:, :2] += 5.0
    y2[:, :, -2:] -= 3.0
    l2, g2 = D.net_grad(p, x, y2, False, margin=2)
    assert l1 == l2 and np.array_equal(g1, g2)
    inner = (out - y)[:, 2:5, 2:6, 2:7]
    assert abs(l1 - np.mean(inner ** 2)) < 1e-12 * l1


def test_projection_restores_continuity():
    rng = np.random.default_rng(6)
    n, step, dt = (8, 6, 10), (1.0, 0.5, 2.0), 0.3
    j = [_rand(rng, *n) for _ in range(3)]
    r0 = _rand(rng, *n)
    r1 = r0 + 0.1 * _rand(rng, *n)
    r1 -= r1.mean() - r0.mean()                       # charge is conserved
    jp = D.project_continuity(*j, r0, r1, dt, step)
    res = D.divergence_periodic(*jp, step) + (r1 - r0) / dt
    scale = np.abs(D.divergence_periodic(*j, step)).max()
    assert np.abs(res).max() < 1e-12 * scale
    again = D.project_continuity(*jp, r0, r1, dt, step)
    assert max(np.abs(a - b).max() for a, b in zip(again, jp)) < 1e-12 * max(np.abs(a).max() for a in jp)
