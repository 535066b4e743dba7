"""TEST INFRASTRUCTURE ONLY: NumPy restatement of the learned warm-start.

Only tests/ and bench.py's CPU legs import this module; it checks the GPU
kernels and is never the thing measured or shipped.

The warm start has no counterpart in the reference. What it restates:
  * the network: 12 -> 64 -> 64 -> 3, GELU (tanh form, Hendrycks & Gimpel
    2016) after both hidden layers, parameters W1 [12][64], b1, W2 [64][64],
    b2, W3 [64][3], b3 (5187 floats);
  * the GPU's arithmetic points: operands of every product rounded to BF16
    (round to nearest even), products accumulated exactly here (FP32 on the
    GPU), biases and GELU unrounded (`emulate=True`); `emulate=False` is the
    plain FP64 network;
  * the loss sum_v w_v |y_v - t_v|^2 / sum_v w_v and its gradient by
    back-propagation, checked against central differences in the tests;
  * Adam (Kingma & Ba 2015; beta1 0.9, beta2 0.999, eps 1e-8);
  * the one-ring features (strains: the errors of Physics/ClothPBD.cpp:78
    warp, :89 weft, :101 shear) and the residual of the solve: the warp and
    weft errors and the bend error s theta of ClothPBD.cpp:179-224 (the
    compliant shear, ClothPBD.cpp:96-101, is reported apart).
"""
import numpy as np

N_IN, N_HID, N_OUT = 12, 64, 3
OFF = dict(W1=0, b1=768, W2=832, b2=4928, W3=4992, b3=5184)
NUM_PARAMS = 5187
SCALE_CORR, SCALE_STRAIN, SCALE_OUT = 10.0, 100.0, 10.0
GELU_C, GELU_A = 0.7978845608, 0.044715


def bf16(x):
    """FP32 values rounded to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.reshape(np.shape(x))


def split(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[0:768].reshape(12, 64), b1=p[768:832], W2=p[832:4928].reshape(64, 64), b2=p[4928:4992],
                W3=p[4992:5184].reshape(64, 3), b3=p[5184:5187])


def join(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(), d["b3"]])


def gelu(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _q(x, emulate):
    return bf16(x).astype(np.float64) if emulate else np.asarray(x, dtype=np.float64)


def forward(p, X, emulate=True):
    """y (n x 3) and the intermediates of the forward pass."""
    P = split(p)
    if emulate:
        W1, W2, W3 = (_q(P[k].astype(np.float32), True) for k in ("W1", "W2", "W3"))
        x = _q(np.asarray(X, dtype=np.float32), True)
    else:
        W1, W2, W3 = P["W1"], P["W2"], P["W3"]
        x = np.asarray(X, dtype=np.float64)
    z1 = x @ W1 + P["b1"]
    h1 = _q(gelu(z1).astype(np.float32), emulate) if emulate else gelu(z1)
    z2 = h1 @ W2 + P["b2"]
    h2 = _q(gelu(z2).astype(np.float32), emulate) if emulate else gelu(z2)
    y = h2 @ W3 + P["b3"]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1, W2=W2, W3=W3)


def loss_and_grad(p, X, T, w=None, emulate=True):
    """Weighted mean squared error and its gradient (parameter order)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    w = np.ones(len(X)) if w is None else np.asarray(w, dtype=np.float64)
    inv = 1.0 / w.sum()
    y, a = forward(p, X, emulate)
    e = y - T
    loss = float((w[:, None] * e * e).sum() * inv)
    dy = w[:, None] * 2.0 * e * inv
    q = (lambda v: _q(v.astype(np.float32), True)) if emulate else (lambda v: v)
    dyq = q(dy)
    g = {"W3": a["h2"].T @ dyq, "b3": dy.sum(0)}
    dz2 = (dyq @ a["W3"].T) * gelu_grad(a["z2"])
    dz2q = q(dz2)
    g["W2"] = a["h1"].T @ dz2q
    g["b2"] = dz2.sum(0)
    dz1 = (dz2q @ a["W2"].T) * gelu_grad(a["z1"])
    dz1q = q(dz1)
    g["W1"] = a["x"].T @ dz1q
    g["b1"] = dz1.sum(0)
    return loss, join(g)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step at step number t (1-based); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------- features

def tri_strains(pos, tri_idx, rest):
    """Per triangle (warp, weft, shear) at pos; rest rows are the oracle's
    (du1 du2 dv1 dv2 det invDet lu0 lv0 area lu2 lv2)."""
    pos = np.asarray(pos, dtype=np.float64)
    x1, x2, x3 = pos[tri_idx[:, 0]], pos[tri_idx[:, 1]], pos[tri_idx[:, 2]]
    du1, du2, dv1, dv2 = rest[:, 0:1], rest[:, 1:2], rest[:, 2:3], rest[:, 3:4]
    invDet, lu2, lv2 = rest[:, 5], rest[:, 9], rest[:, 10]
    dx1, dx2 = x2 - x1, x3 - x1
    wu = invDet[:, None] * (dv2 * dx1 - dv1 * dx2)
    wv = invDet[:, None] * (-du2 * dx1 + du1 * dx2)
    return np.stack([0.5 * ((wu * wu).sum(1) - lu2), 0.5 * ((wv * wv).sum(1) - lv2), (wu * wv).sum(1)], axis=1)


def one_ring(nv, edges):
    """Sorted neighbour lists from (i1, i2, ...) edge rows."""
    nb = [set() for _ in range(nv)]
    for e in edges:
        nb[int(e[0])].add(int(e[1]))
        nb[int(e[1])].add(int(e[0]))
    return [sorted(s) for s in nb]


def features(xhat, prev, corr, tri_idx, rest, neighbours, ell):
    """The 12 features per vertex (before BF16 rounding)."""
    xhat = np.asarray(xhat, dtype=np.float64)
    corr = np.asarray(corr, dtype=np.float64)
    nv = len(xhat)
    X = np.zeros((nv, 12))
    X[:, 0:3] = (xhat - np.asarray(prev, dtype=np.float64)) / ell
    X[:, 3:6] = SCALE_CORR * corr / ell
    for v in range(nv):
        if neighbours[v]:
            X[v, 6:9] = SCALE_CORR * (corr[neighbours[v]].mean(0) - corr[v]) / ell
    S = tri_strains(xhat, tri_idx, rest)
    acc = np.zeros((nv, 3))
    cnt = np.zeros(nv)
    for k in range(3):
        np.add.at(acc, tri_idx[:, k], S)
        np.add.at(cnt, tri_idx[:, k], 1)
    X[:, 9:12] = SCALE_STRAIN * acc / np.maximum(cnt, 1)[:, None]
    return X


# ---------------------------------------------------------------- residual

def bend_angles(pos, bends):
    pos = np.asarray(pos, dtype=np.float64)
    x1, x2, x3, x4 = (pos[bends[:, k]] for k in range(4))
    x12, x13, x14 = x2 - x1, x3 - x1, x4 - x1
    n1 = np.cross(x12, x13)
    n2 = np.cross(x14, x12)
    l1 = np.linalg.norm(n1, axis=1)
    l2 = np.linalg.norm(n2, axis=1)
    ok = (l1 > 0) & (l2 > 0)
    c = np.zeros(len(bends))
    c[ok] = (n1[ok] * n2[ok]).sum(1) / (l1[ok] * l2[ok])
    theta = np.zeros(len(bends))
    theta[ok] = np.arccos(np.clip(c[ok], -1.0, 1.0))
    return theta


def residual_terms(pos, tri_idx, rest, bends, s_stiff):
    """RMS of (warp and weft together, shear, s theta)."""
    S = tri_strains(pos, tri_idx, rest)
    th = s_stiff * bend_angles(pos, bends)
    return (float(np.sqrt((S[:, :2] ** 2).sum() / (2 * len(tri_idx)))), float(np.sqrt((S[:, 2] ** 2).mean())),
            float(np.sqrt((th * th).mean())) if len(th) else 0.0)


def residual(pos, tri_idx, rest, bends, s_stiff):
    """RMS residual of the stiff constraints: sqrt((sum warp^2 + weft^2 +
    sum (s theta)^2) / (2T + B)); the compliant shear is left out."""
    S = tri_strains(pos, tri_idx, rest)
    th = s_stiff * bend_angles(pos, bends)
    return float(np.sqrt(((S[:, :2] ** 2).sum() + (th * th).sum()) / (2 * len(tri_idx) + len(bends))))
