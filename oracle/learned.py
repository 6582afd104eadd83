"""CPU restatement of the learned troubled-face indicator (numpy, float64 unless stated).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use it, as the checker.
The product (fenice_amd) never imports it.

The indicator has no reference counterpart beyond the functions its labels come from:
- label: the departure of the WENO5-Z weights from the linear weights (0.1, 0.6, 0.3) (weno5z, Schemes.cpp:338-353),
  max over the primitive variables and both face states of a face, max_k |w_k - d_k| with
  w_k = a_k / (a0 + a1 + a2); a face is troubled when it exceeds the label threshold. Density and pressure are
  regularised as WENO5-Z regularises them (eps_rel times their mean square); velocity components by eps_rel times
  the face's mean |v|^2 + p / rho over its six cells, so that rounding noise in a component that is locally near
  zero does not count as a jump;
- a second label source: where the MUSCL limiter clips, i.e. how far the limited slope phi(a, b) of a face value
  (apply_limiter, Schemes.hpp:110-117) lies from the central slope (a + b) / 2, relative to the differences'
  magnitude (clip); the label of a face may be the departure, the clip, or the larger of the two;
- smooth faces take the linear upwind-5 value (2c0 - 13c1 + 47c2 + 27c3 - 3c4) / 60, which is weno5z with a = d
  (exact for quartic cell-average data);
- features of a face (15): for density, pressure and the velocity component with the largest total variation over
  the face's six-cell stencil c0..c5 (cells t-3 .. t+2 of a face between cells t-1 and t), the four second
  differences d_{j+1} - d_j and the largest |d_j| of its first differences d_j = c_{j+1} - c_j, divided by
  sqrt(sum of d_j^2 + eps_rel * mean(c^2) + 1e-30), eps_rel the WENO regularisation (the same relative scaling
  WENO5-Z applies);
- network: 15 -> 32 (ReLU) -> 32 (ReLU) -> 1 logit; troubled when logit > threshold. Inference rounds x, W1, b1, W2
  and the first hidden layer to bfloat16 (round to nearest even) and accumulates in float32.
Parameter vector layout (float32, 1601 values): W1 (32 x 15, row-major), b1 (32), W2 (32 x 32), b2 (32), w3 (32),
b3 (1).
"""
from __future__ import annotations

import numpy as np

NIN, NH = 15, 32
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1
D_LIN = np.array([0.1, 0.6, 0.3])


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float32)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


# ---- stencils on a periodic mesh ------------------------------------------------------------------------------
def stencils(W, axis):
    """Six-cell stencils of every face normal to `axis` of a periodic mesh. W: primitive planes (nv, nz, ny, nx)
    (2-D: nz = 1). Returns (6, nv, nfaces) with faces in the device's per-axis order; on a periodic axis the face
    t = n is the face t = 0, so the n faces t = 0 .. n-1 are returned."""
    ax = {0: 3, 1: 2, 2: 1}[axis]
    cells = [np.roll(W, 3 - s, axis=ax) for s in range(6)]   # cell t - 3 + s at position t
    return np.stack([c.reshape(W.shape[0], -1) for c in cells])


def weno5z_weights(c0, c1, c2, c3, c4, eps_rel, m=None):
    """WENO5-Z weights with the regularisation eps_rel * m (m: the mean square of the five values by default)."""
    if m is None:
        m = (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4) * 0.2
    eps = eps_rel * m + 1e-300
    s0, t0 = c0 - 2.0 * c1 + c2, c0 - 4.0 * c1 + 3.0 * c2
    s1, t1 = c1 - 2.0 * c2 + c3, c1 - c3
    s2, t2 = c2 - 2.0 * c3 + c4, 3.0 * c2 - 4.0 * c3 + c4
    b0 = (13.0 / 12.0) * s0 * s0 + 0.25 * t0 * t0
    b1 = (13.0 / 12.0) * s1 * s1 + 0.25 * t1 * t1
    b2 = (13.0 / 12.0) * s2 * s2 + 0.25 * t2 * t2
    tau = np.abs(b0 - b2)
    a = np.stack([0.1 * (1.0 + tau / (b0 + eps)), 0.6 * (1.0 + tau / (b1 + eps)), 0.3 * (1.0 + tau / (b2 + eps))])
    return a / a.sum(axis=0)


def velocity_scale(st, nvel):
    """mean over the six cells of |v|^2 + p / rho: the regularisation scale of the velocity components. st: (6, nv, nf)."""
    vel = [1, 2] if nvel == 2 else [1, 2, 4]
    v2 = sum(st[:, q] ** 2 for q in vel)
    return (v2 + st[:, 3] / st[:, 0]).mean(axis=0)


def departure(st, eps_rel, nvel):
    """Label value of each face: max over variables and both face states of max_k |w_k - d_k|. st: (6, nv, nf).
    Density and pressure are regularised by their own mean square (as WENO5-Z is), velocity components by the
    face's kinetic-plus-acoustic scale (velocity_scale), so rounding noise in a near-zero component is no jump."""
    c = st
    vs = velocity_scale(st, nvel)
    vel = [1, 2] if nvel == 2 else [1, 2, 4]
    out = np.zeros(st.shape[2])
    for q in range(st.shape[1]):
        m = vs if q in vel else None
        wl = weno5z_weights(c[0, q], c[1, q], c[2, q], c[3, q], c[4, q], eps_rel, m)
        wr = weno5z_weights(c[5, q], c[4, q], c[3, q], c[2, q], c[1, q], eps_rel, m)
        d = np.maximum(np.abs(wl - D_LIN[:, None]).max(axis=0), np.abs(wr - D_LIN[:, None]).max(axis=0))
        out = np.maximum(out, d)
    return out


def limiter(lim, a, b):
    """The reference's slope limiters phi(a, b) (Schemes.hpp:82-117): 0 minmod, 1 van Leer, 2 MC, 3 Koren."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    same = a * b > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        if lim == 0:
            return 0.5 * (np.sign(a) + np.sign(b)) * np.minimum(np.abs(a), np.abs(b))
        if lim == 1:
            return np.where(same, 2.0 * a * b / (a + b), 0.0)
        if lim == 2:
            return np.where(same, np.minimum(np.minimum(2.0 * np.abs(a), 0.5 * np.abs(a + b)), 2.0 * np.abs(b)) * np.sign(a), 0.0)
        return np.where(same, np.minimum(np.minimum(2.0 * np.abs(a), np.abs(2.0 * a + b) / 3.0), 2.0 * np.abs(b)) * np.sign(a), 0.0)


def clip(st, eps_rel, nvel, lim):
    """MUSCL-limiter clip label value of each face: max over the variables and both face values of
    |phi(a, b) - (a + b) / 2| / sqrt((a^2 + b^2) / 2 + eps_rel * m + 1e-30), on the four MUSCL cells c1..c4 of the
    six-cell stencil; a = c2 - c1, b = c3 - c2 for the left value, a = c4 - c3, b = c3 - c2 for the right one (the
    reconstruction's wl = c1 + phi(c1 - c0, c2 - c1) / 2, wr = c2 - phi(c3 - c2, c2 - c1) / 2, Schemes.cpp:261-262,
    on cells t-2 .. t+1). m: the mean square of the four values (density, pressure) or the face's velocity scale."""
    vs = velocity_scale(st, nvel)
    vel = [1, 2] if nvel == 2 else [1, 2, 4]
    out = np.zeros(st.shape[2])
    for q in range(st.shape[1]):
        c1, c2, c3, c4 = st[1, q], st[2, q], st[3, q], st[4, q]
        m = vs if q in vel else (c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4) * 0.25
        b = c3 - c2
        for a in (c2 - c1, c4 - c3):
            phi = limiter(lim, a, b)
            out = np.maximum(out, np.abs(phi - 0.5 * (a + b)) / np.sqrt(0.5 * (a * a + b * b) + eps_rel * m + 1e-30))
    return out


def label_value(st, eps_rel, nvel, source="weno5", lim=1):
    """The label value of each face: the WENO5-Z departure, the MUSCL clip, or the larger of the two ("both")."""
    if source == "weno5":
        return departure(st, eps_rel, nvel)
    if source == "muscl":
        return clip(st, eps_rel, nvel, lim)
    return np.maximum(departure(st, eps_rel, nvel), clip(st, eps_rel, nvel, lim))


def upwind5(c0, c1, c2, c3, c4):
    return (2.0 * c0 - 13.0 * c1 + 47.0 * c2 + 27.0 * c3 - 3.0 * c4) / 60.0


def features(st, eps_rel, nvel):
    """(nf, 15) features of each face from its stencils (6, nv, nf); velocity slots are 1, 2 (and 4 in 3-D)."""
    vel = [1, 2] if nvel == 2 else [1, 2, 4]
    tv = np.stack([np.abs(np.diff(st[:, q, :], axis=0)).sum(axis=0) for q in vel])
    pick = np.argmax(tv, axis=0)   # first maximum on ties
    v = np.choose(pick, [st[:, q, :] for q in vel])
    out = []
    for c in (st[:, 0, :], st[:, 3, :], v):
        d = np.diff(c, axis=0)
        norm = np.sqrt((d * d).sum(axis=0) + eps_rel * (c * c).mean(axis=0) + 1e-30)
        out.append((np.diff(d, axis=0) / norm).T)
        out.append((np.abs(d).max(axis=0) / norm)[:, None])
    return np.concatenate(out, axis=1)


# ---- network --------------------------------------------------------------------------------------------------
def forward(X, p, bf16_inference=False):
    """Logits of the faces X (n, 15)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if bf16_inference:   # b1 rides in the bfloat16 weight column of the constant-1 input
        X = bf16(X).astype(np.float64)
        W1 = bf16(W1).astype(np.float64)
        W2 = bf16(W2).astype(np.float64)
        h1 = np.maximum(X @ W1.T + bf16(b1).astype(np.float64), 0.0)
        h1 = bf16(h1).astype(np.float64)
    else:
        h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss_and_grad(X, y, p, pos_weight=1.0):
    """Mean weighted binary cross-entropy with logits, and its gradient with respect to the parameter vector
    (float64 backpropagation)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = X.shape[0]
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    z = h2 @ w3 + b3
    w = np.where(y > 0.5, pos_weight, 1.0)
    # stable log(1 + exp(-|z|)) form
    lse = np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))
    loss = (w * lse).sum() / n
    sig = 1.0 / (1.0 + np.exp(-z))
    gz = w * (sig - y) / n
    g_w3 = h2.T @ gz
    g_b3 = gz.sum()
    gh2 = np.outer(gz, w3) * (z2 > 0)
    g_W2 = gh2.T @ h1
    g_b2 = gh2.sum(axis=0)
    gh1 = (gh2 @ W2) * (z1 > 0)
    g_W1 = gh1.T @ X
    g_b1 = gh1.sum(axis=0)
    return loss, np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
