"""CPU oracle of the learned form: the current denoiser and the continuity projection.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU leg as
the checker; the library never calls it.

pica has no counterpart (no learned or projected current).  What pins these functions:
  - conv3d_periodic / net_forward: a direct sum over the 27 taps in FP64 (the definition of a
    periodic 3x3x3 cross-correlation); tests pin it against a brute-force single-voxel sum and
    shift equivariance.
  - net_grad: exact FP64 backward of net_forward; tests pin it with central finite differences.
  - project_continuity: spectral solve of the 7-point periodic Laplacian; tests pin it by the
    property it exists for (discrete continuity at round-off after projection) and by
    idempotence.

Layout conventions shared with the device code (pica_amd/csrc/pk_denoise.hip):
  - a volume is [C][n0][n1][n2] (z fastest), periodic in every dimension;
  - tap t = (dx+1)*9 + (dy+1)*3 + (dz+1) reads the input at (x+dx, y+dy, z+dz);
  - parameters, FP32: W1[27][3][16], b1[16], W2[27][16][16], b2[16], W3[27][16][3], b3[3];
  - BF16 rounding (round to nearest even) at the points the device rounds: the input, the
    weights, the ReLU outputs a1 and a2, and the back-propagated gradients g3, dz2, dz1.
"""
import numpy as np

HIDDEN = 16
SHAPES = [("W1", (27, 3, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (27, HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (27, HIDDEN, 3)), ("b3", (3,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def taps():
    """(dx, dy, dz) of tap t"""
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16(x):
    """round FP32 values to the nearest BF16 (ties to even), returned as float64"""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(a.shape)


def unpack(p):
    """parameter vector -> dict of arrays (FP64 copies)"""
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def shifted(x, dx, dy, dz):
    """y[..., i, j, k] = x[..., i+dx, j+dy, k+dz] (periodic)"""
    return np.roll(x, (-dx, -dy, -dz), axis=(-3, -2, -1))


def conv3d_periodic(x, W, b=None):
    """out[co, v] = b[co] + sum_t sum_ci W[t, ci, co] x[ci, v + off(t)], FP64"""
    cin, cout = W.shape[1], W.shape[2]
    assert x.shape[0] == cin
    out = np.zeros((cout,) + x.shape[1:])
    for t, (dx, dy, dz) in enumerate(taps()):
        xs = shifted(x, dx, dy, dz).reshape(cin, -1)
        out += (W[t].T @ xs).reshape(out.shape)
    if b is not None:
        out += np.asarray(b, np.float64)[:, None, None, None]
    return out


def conv3d_transpose(g, W):
    """adjoint of conv3d_periodic in its input: dx[ci, v] = sum_t sum_co W[t, ci, co] g[co, v - off(t)]"""
    cin = W.shape[1]
    out = np.zeros((cin,) + g.shape[1:])
    for t, (dx, dy, dz) in enumerate(taps()):
        gs = shifted(g, -dx, -dy, -dz).reshape(g.shape[0], -1)
        out += (W[t] @ gs).reshape(out.shape)
    return out


def weight_grad(x, g):
    """dW[t, ci, co] = sum_v x[ci, v + off(t)] g[co, v]"""
    cin, cout = x.shape[0], g.shape[0]
    dW = np.zeros((27, cin, cout))
    gf = g.reshape(cout, -1)
    for t, (dx, dy, dz) in enumerate(taps()):
        dW[t] = shifted(x, dx, dy, dz).reshape(cin, -1) @ gf.T
    return dW


def net_forward(p, x, emulate=True):
    """the 3 -> 16 -> 16 -> 3 network; emulate=True rounds to BF16 where the device does.
    Returns (out, cache) with out FP64 [3][n0][n1][n2]."""
    P = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W = {k: r(P[k]) if k[0] == "W" else P[k] for k in P}
    xb = r(x)
    z1 = conv3d_periodic(xb, W["W1"], W["b1"])
    a1 = r(np.maximum(z1, 0.0))
    z2 = conv3d_periodic(a1, W["W2"], W["b2"])
    a2 = r(np.maximum(z2, 0.0))
    out = conv3d_periodic(a2, W["W3"], W["b3"])
    return out, {"W": W, "xb": xb, "a1": a1, "a2": a2, "z1": z1, "z2": z2, "r": r}


def interior(shape, margin):
    """mask [n0][n1][n2] of the voxels at least `margin` from every face (all of them for margin 0)"""
    m = np.zeros(shape, bool)
    m[margin:shape[0] - margin, margin:shape[1] - margin, margin:shape[2] - margin] = True
    return m


def net_loss(p, x, y, emulate=True, margin=0):
    """mean squared error over the channels of the interior voxels (see `interior`)"""
    out, _ = net_forward(p, x, emulate)
    m = interior(out.shape[1:], margin)
    return float(np.sum(((out - y) ** 2)[:, m]) / (3 * m.sum()))


def net_grad(p, x, y, emulate=True, margin=0):
    """(loss, gradient vector) of the mean squared error over the interior voxels; with emulate=True
    the back-propagated gradients are rounded to BF16 where the device rounds them"""
    out, c = net_forward(p, x, emulate)
    r, W = c["r"], c["W"]
    m = interior(out.shape[1:], margin)
    diff = (out - y) * m
    count = 3 * m.sum()
    loss = float(np.sum(diff ** 2) / count)
    g3 = r(2.0 * diff / count)
    G = {"W3": weight_grad(c["a2"], g3), "b3": g3.reshape(3, -1).sum(1)}
    dz2 = r(conv3d_transpose(g3, W["W3"]) * (c["a2"] > 0))
    G["W2"] = weight_grad(c["a1"], dz2)
    G["b2"] = dz2.reshape(HIDDEN, -1).sum(1)
    dz1 = r(conv3d_transpose(dz2, W["W2"]) * (c["a1"] > 0))
    G["W1"] = weight_grad(c["xb"], dz1)
    G["b1"] = dz1.reshape(HIDDEN, -1).sum(1)
    return loss, pack(G)


def init_params(seed):
    """He-normal weights, zero biases (FP32 values)"""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name[0] == "W":
            d[name] = rng.standard_normal(shp) * np.sqrt(2.0 / (27 * shp[1]))
        else:
            d[name] = np.zeros(shp)
    return pack(d).astype(np.float32)


# ---------------------------------------------------------------- continuity projection
def divergence_periodic(jx, jy, jz, step):
    """node i: sum_d (J_d[i + e_d] - J_d[i]) / h_d on the periodic physical cells"""
    return ((np.roll(jx, -1, 0) - jx) / step[0] + (np.roll(jy, -1, 1) - jy) / step[1] +
            (np.roll(jz, -1, 2) - jz) / step[2])


def project_continuity(jx, jy, jz, rho_prev, rho_now, dt, step):
    """J - grad(phi) with lap(phi) = div J + (rho_now - rho_prev) / dt, where grad is the backward
    difference (minus the adjoint of the forward-difference divergence) and lap = div grad is the
    7-point periodic Laplacian, solved exactly in Fourier space (mean of phi = 0)"""
    r = divergence_periodic(jx, jy, jz, step) + (rho_now - rho_prev) / dt
    n = r.shape
    lam = np.zeros(n)
    for d in range(3):
        k = np.arange(n[d])
        ld = -4.0 * np.sin(np.pi * k / n[d]) ** 2 / step[d] ** 2
        sh = [1, 1, 1]
        sh[d] = n[d]
        lam = lam + ld.reshape(sh)
    rh = np.fft.fftn(r)
    lam[0, 0, 0] = 1.0
    ph = rh / lam
    ph[0, 0, 0] = 0.0
    phi = np.real(np.fft.ifftn(ph))
    out = []
    for d, j in enumerate((jx, jy, jz)):
        out.append(j - (phi - np.roll(phi, 1, d)) / step[d])
    return out
