"""CPU restatement of the learned shock sensor of the hybrid reconstruction.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
sensor (nuc3d_amd/csrc/nuc3d_sensor.h, nuc3d_sensor.hip), never by the
product.  The sensor has no counterpart in the reference; what it restates is
this build's own definition:

  features  rho, u, v, w, p on the six cells m-3..m+2 of face m, each centred
            on its stencil mean and divided by (range + 0.01 S), with the
            acoustic scales S = (rho U / c, U, U, U, rho c U) of the stencil's
            mean rho and p, c = sqrt(gamma p / rho) and U = max |velocity
            component| + 0.1 c; FP32 on the FP32 roundings of the
            primitives, then bf16 (round to nearest even);
  teacher   1 when, for some primitive, the eighth-order central face value
            and the WENO5-JS value (weno5js.cpp:75-108 order) of either upwind
            side differ by more than tau S (the same scales in FP64);
  network   30 -> 32 -> 32 -> 1, ReLU, bf16 weights and activations (both
            hidden layers), FP32 accumulation with the biases as initial
            values; loss wt (softplus(z) - y z).

Every step of the features (FP32) and the teacher (FP64) is written in the
same operation order as the device code, so features and labels are
bit-exact.
The network is evaluated here in float64 from the same bf16-rounded operands
(the GPU accumulates in FP32 on the matrix cores), so logits and gradients
agree to a tolerance, not bit for bit.  quantize=False gives the plain
float64 network, whose gradient the tests pin against torch autograd and
finite differences.
"""
from __future__ import annotations

import numpy as np

NFEAT = 30
NP = 2145
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 1024, 1056, 2080, 2112, 2144
FEAT_EPS = 0.01


# ---------------------------------------------------------------- bf16

def bf16_bits(x) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even (finite inputs)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return (u >> 16).astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    """bf16 bit patterns -> their float32 values."""
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def bf16_round(x) -> np.ndarray:
    return bf16_value(bf16_bits(x))


# ---------------------------------------------------------------- stencils

def con2prim(q, jac, gamma):
    """rho, u, v, w, p of conserved/J (physicsModel.cpp:313-353 order, as the
    device's con2prim)."""
    rho = q[0] * jac
    u = q[1] / rho * jac
    v = q[2] / rho * jac
    w = q[3] / rho * jac
    E = q[4] * jac
    p = (E - 0.5 * rho * (u * u + v * v + w * w)) * (gamma - 1.0)
    return np.stack([rho, u, v, w, p])


def weno5js(f0, f1, f2, f3, f4):
    """WENO5-JS face value, weno5js.cpp:75-108 (squares as products)."""
    G0, G1, EPS = 13.0 / 12.0, 1.0 / 4.0, 1.0e-6
    t = f0 - 2.0 * f1 + f2
    u = f0 - 4.0 * f1 + 3.0 * f2
    b0 = G0 * (t * t) + G1 * (u * u)
    t = f1 - 2.0 * f2 + f3
    u = f1 - f3
    b1 = G0 * (t * t) + G1 * (u * u)
    t = f2 - 2.0 * f3 + f4
    u = 3.0 * f2 - 4.0 * f3 + f4
    b2 = G0 * (t * t) + G1 * (u * u)
    q0 = 1.0 / 3.0 * f0 + -7.0 / 6.0 * f1 + 11.0 / 6.0 * f2
    q1 = -1.0 / 6.0 * f1 + 5.0 / 6.0 * f2 + 1.0 / 3.0 * f3
    q2 = 1.0 / 3.0 * f2 + 5.0 / 6.0 * f3 + -1.0 / 6.0 * f4
    e = EPS + b0
    a0 = 0.1 / (e * e)
    e = EPS + b1
    a1 = 0.6 / (e * e)
    e = EPS + b2
    a2 = 0.3 / (e * e)
    s = a0 + a1 + a2
    o0, o1, o2 = a0 / s, a1 / s, a2 / s
    return o0 * q0 + o1 * q1 + o2 * q2


def central8(c0, c1, c2, c3, c4, c5, c6, c7):
    """Eighth-order central face value (FP64), the division by 840 taken as a
    product with its FP64 reciprocal."""
    return ((533.0 * (c3 + c4) - 139.0 * (c2 + c5)) + (29.0 * (c1 + c6) - 3.0 * (c0 + c7))) * (1.0 / 840.0)


_F = np.float32


def central8f(c0, c1, c2, c3, c4, c5, c6, c7):
    """The same in FP32 (the positivity check of the hybrid)."""
    return ((_F(533.0) * (c3 + c4) - _F(139.0) * (c2 + c5)) + (_F(29.0) * (c1 + c6) - _F(3.0) * (c0 + c7))) * \
        (_F(1.0) / _F(840.0))


def stencil_scales(P6, gamma):
    """FP64 scales S of the teacher over the stencil P6[v][j] (v = rho,u,v,w,p;
    j = cells m-3..m+2) and the validity mask: mean rho and p, c = sqrt(gamma
    mean p / mean rho), U = max |velocity component| + 0.1 c, and
    S = (rho U / c, U, U, U, rho c U)."""
    mean = {}
    for v in (0, 4):
        s = P6[v][0]
        for j in range(1, 6):
            s = s + P6[v][j]
        mean[v] = s / 6.0
    rb, pb = mean[0], mean[4]
    ok = (rb > 0.0) & (pb > 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        cb = np.sqrt(gamma * pb / rb)
        um = np.zeros(np.shape(rb))
        for v in (1, 2, 3):
            for j in range(6):
                um = np.maximum(um, np.abs(P6[v][j]))
        U = um + 0.1 * cb
        S = [rb * U / cb, U, U, U, rb * cb * U]
    return S, ok


def features(P6, gamma):
    """bf16 bit patterns (..., 32) of the features of stencils P6 (5, 6, ...)
    of FP64 primitives, and the validity mask; invalid stencils get zero
    features.  FP32 arithmetic on the FP32 roundings of the primitives:
    mean = sum * (1/6), c = sqrt(gamma mean p / mean rho),
    U = max |velocity component| + 0.1 c, S = (rho U / c, U, U, U, rho c U),
    x = (P - mean) * (1 / (range + 0.01 S))."""
    P = [[np.asarray(P6[v][j], dtype=np.float64).astype(_F) for j in range(6)] for v in range(5)]
    g = _F(gamma)
    sixth = _F(1.0) / _F(6.0)

    def mean_of(v):
        s = P[v][0]
        for j in range(1, 6):
            s = s + P[v][j]
        return s * sixth

    rb, pb = mean_of(0), mean_of(4)
    ok = (rb > 0) & (pb > 0)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        cb = np.sqrt(g * pb / rb)
        um = np.zeros(np.shape(rb), dtype=_F)
        for v in (1, 2, 3):
            for j in range(6):
                um = np.maximum(um, np.abs(P[v][j]))
        U = um + _F(0.1) * cb
        S = [rb * U / cb, U, U, U, rb * cb * U]
        shape = np.shape(P[0][0])
        x = np.zeros(shape + (32,), dtype=np.uint16)
        for v in range(5):
            mean = mean_of(v)
            lo, hi = P[v][0], P[v][0]
            for j in range(1, 6):
                lo = np.minimum(lo, P[v][j])
                hi = np.maximum(hi, P[v][j])
            inv = _F(1.0) / ((hi - lo) + _F(0.01) * S[v])
            for j in range(6):
                val = (P[v][j] - mean) * inv
                x[..., 6 * v + j] = np.where(ok, bf16_bits(np.where(ok, val, _F(0.0))), 0)
    return x, ok


def teacher(P8, gamma, tau):
    """Labels of stencils P8 (5, 8, ...) over cells m-4..m+3: 1 / 0, -1 where
    the stencil is invalid."""
    S, ok = stencil_scales([P8[v][1:7] for v in range(5)], gamma)
    dev = np.zeros(np.shape(P8[0][0]))
    with np.errstate(invalid="ignore", divide="ignore"):
        for v in range(5):
            c = P8[v]
            c8 = central8(*[c[j] for j in range(8)])
            wp = weno5js(c[1], c[2], c[3], c[4], c[5])
            wm = weno5js(c[6], c[5], c[4], c[3], c[2])
            d = np.maximum(np.abs(c8 - wp), np.abs(c8 - wm)) / S[v]
            dev = np.maximum(dev, d)
    return np.where(ok, (dev > tau).astype(np.int64), -1)


# rotated face-array axes of direction d for arrays of shape (nz, ny, nx):
# (a3, a2, a1) with a1 along d, a2 = axis d+1, a3 = axis d+2
_ROT = {0: (0, 1, 2), 1: (2, 0, 1), 2: (1, 2, 0)}


def rotate(a, d):
    """(..., nz, ny, nx) -> (..., n_{d+2}, n_{d+1}, n_d)."""
    lead = a.ndim - 3
    return np.transpose(a, tuple(range(lead)) + tuple(lead + t for t in _ROT[d]))


def sample_block(q, jac, gamma, tau, d, periodic, stride=1, offset=0, ghosts=3):
    """Samples of direction d of one block (nuc3d_sensor_sample): q (5, nz,
    ny, nx) conserved/J, jac (nz, ny, nx).  periodic: the block's two faces
    of direction d are joined to each other (wrap ghosts, `ghosts` layers
    held); otherwise they are physical faces.  Returns (x bits (n, 32),
    labels (n,)) for the faces of the rotated face array with
    L % stride == offset."""
    prim = con2prim(q, jac, gamma)            # (5, nz, ny, nx)
    pr = rotate(prim, d)                      # (5, a3, a2, n1)
    n1 = pr.shape[-1]
    if periodic:
        ext = np.concatenate([pr[..., n1 - 4:], pr, pr[..., :4]], axis=-1)
        lo, hi = 4 - ghosts, n1 + ghosts - 4   # the eight cells must lie in the ghost frame
    else:
        ext = np.concatenate([np.zeros(pr.shape[:-1] + (4,)), pr, np.zeros(pr.shape[:-1] + (4,))], axis=-1)
        lo, hi = 4, n1 - 4
    m = np.arange(n1 + 1)
    P8 = [[ext[v][..., m + j] for j in range(8)] for v in range(5)]   # cells m-4+j at ext index m+j
    x, ok = features([P8[v][1:7] for v in range(5)], gamma)
    y = teacher(P8, gamma, tau)
    inrange = (m >= lo) & (m <= hi)
    valid = ok & inrange
    y = np.where(valid, y, -1)
    x = np.where(valid[..., None], x, 0).astype(np.uint16)
    x = x.reshape(-1, 32)
    y = y.reshape(-1)
    L = np.arange(y.size)
    sel = (L % stride) == offset
    return x[sel], y[sel].astype(np.float32)


def face_stencil_prims(q, jac, gamma, d, periodic):
    """P6 and P8 stencils of every face of direction d (rotated face array)."""
    prim = con2prim(q, jac, gamma)
    pr = rotate(prim, d)
    n1 = pr.shape[-1]
    if periodic:
        ext = np.concatenate([pr[..., n1 - 4:], pr, pr[..., :4]], axis=-1)
    else:
        ext = np.concatenate([np.zeros(pr.shape[:-1] + (4,)), pr, np.zeros(pr.shape[:-1] + (4,))], axis=-1)
    m = np.arange(n1 + 1)
    return [[ext[v][..., m + j] for j in range(8)] for v in range(5)]


def decisions(q, jac, gamma, d, periodic, params, thr):
    """Hybrid decisions of every face of direction d (0 central, 1 WENO5-JS by
    the network, 2 WENO5-JS by the fallback, 3 physical-face closure) and the
    logits, for a block whose faces of direction d are periodic or physical
    (width 4)."""
    P8 = face_stencil_prims(q, jac, gamma, d, periodic)
    x, ok = features([P8[v][1:7] for v in range(5)], gamma)
    with np.errstate(invalid="ignore"):
        r32 = [np.asarray(c, dtype=np.float64).astype(_F) for c in P8[0]]
        p32 = [np.asarray(c, dtype=np.float64).astype(_F) for c in P8[4]]
        posok = ok & (central8f(*r32) > 0) & (central8f(*p32) > 0)
    z, _ = forward(x.reshape(-1, 32), params)
    z = z.reshape(x.shape[:-1])
    flag = np.where(~posok, 2, np.where(z >= thr, 1, 0))
    if not periodic:
        n1 = flag.shape[-1] - 1
        m = np.arange(n1 + 1)
        flag = np.where((m < 4) | (m > n1 - 4), 3, flag)
    return flag.astype(np.uint8), z


# ---------------------------------------------------------------- network

def unpack(params):
    p = np.asarray(params, dtype=np.float32)
    W1 = p[OW1:OB1].reshape(32, 32)
    b1 = p[OB1:OW2]
    W2 = p[OW2:OB2].reshape(32, 32)
    b2 = p[OB2:OW3]
    w3 = p[OW3:OB3]
    b3 = p[OB3]
    return W1, b1, W2, b2, w3, b3


def init_params(seed=0):
    """He-initialised parameters (FP32), zero biases, zero padding inputs."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, dtype=np.float32)
    W1 = rng.normal(0.0, np.sqrt(2.0 / NFEAT), (32, 32)).astype(np.float32)
    W1[:, NFEAT:] = 0.0
    p[OW1:OB1] = W1.ravel()
    p[OW2:OB2] = rng.normal(0.0, np.sqrt(2.0 / 32), (32, 32)).astype(np.float32).ravel()
    p[OW3:OB3] = rng.normal(0.0, np.sqrt(1.0 / 32), 32).astype(np.float32)
    return p


def _q(a, quantize):
    return bf16_round(a).astype(np.float64) if quantize else np.asarray(a, dtype=np.float64)


def forward(xbits, params, quantize=True):
    """Logits of samples whose features are bf16 bit patterns (n, 32) (or
    float features (n, 32) when quantize=False).  Returns (z, cache)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    x = bf16_value(xbits).astype(np.float64) if quantize else np.asarray(xbits, dtype=np.float64)
    W1q, W2q, w3q = _q(W1, quantize), _q(W2, quantize), _q(w3, quantize)
    if quantize:
        W1q[:, NFEAT:] = 0.0
    h1 = x @ W1q.T + b1.astype(np.float64)
    a1 = _q(np.maximum(h1, 0.0).astype(np.float32), quantize) if quantize else np.maximum(h1, 0.0)
    h2 = a1 @ W2q.T + b2.astype(np.float64)
    a2 = _q(np.maximum(h2, 0.0).astype(np.float32), quantize) if quantize else np.maximum(h2, 0.0)
    z = a2 @ w3q + float(b3)
    return z, (x, h1, a1, h2, a2, W2q, w3q)


def loss_grad(xbits, y, params, pos_weight=1.0, quantize=True, batch=None):
    """Mean loss over the batch and its gradient (NP values), as the training
    kernels define them: wt (softplus(z) - y z), wt = pos_weight (y = 1),
    1 (y = 0), 0 (y < 0), divided by the batch size."""
    y = np.asarray(y, dtype=np.float64)
    n = batch or y.size
    z, (x, h1, a1, h2, a2, W2q, w3q) = forward(xbits, params, quantize)
    wt = np.where(y < 0, 0.0, np.where(y > 0.5, pos_weight, 1.0)) / n
    yy = (y > 0.5).astype(np.float64)
    sig = 1.0 / (1.0 + np.exp(-z))
    sp = np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))
    loss = float(np.sum(wt * (sp - yy * z)))
    dz = wt * (sig - yy)
    dh2 = np.where(h2 > 0.0, dz[:, None] * w3q[None, :], 0.0)
    dh2q = _q(dh2.astype(np.float32), quantize) if quantize else dh2
    dA1 = dh2q @ W2q
    dh1 = np.where(h1 > 0.0, dA1, 0.0)
    dh1q = _q(dh1.astype(np.float32), quantize) if quantize else dh1
    g = np.zeros(NP)
    gW1 = dh1q.T @ x
    gW1[:, NFEAT:] = 0.0
    g[OW1:OB1] = gW1.ravel()
    g[OB1:OW2] = dh1.sum(axis=0)
    g[OW2:OB2] = (dh2q.T @ a1).ravel()
    g[OB2:OW3] = dh2.sum(axis=0)
    g[OW3:OB3] = (dz[:, None] * a2).sum(axis=0)
    g[OB3] = dz.sum()
    return loss, g


_M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def batch_indices(seed, step, batch, count):
    """Dataset rows of training step `step` (counted from 0 over the sensor's
    life): row of batch position p = mix(seed ^ mix(step * 0x100000001B3 + p))
    mod count, as k_sensor_fwdbwd draws them."""
    return np.array([_mix64((seed ^ _mix64((step * 0x100000001B3 + p) & _M64)) & _M64) % count
                     for p in range(batch)], dtype=np.int64)


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); the padding inputs of W1 stay fixed."""
    params = np.array(params, dtype=np.float64)
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    step = lr * mh / (np.sqrt(vh) + eps)
    mask = np.ones(NP, dtype=bool)
    mask[OW1:OB1] = (np.arange(1024) % 32) < NFEAT
    params[mask] -= step[mask]
    return params.astype(np.float32), m, v
