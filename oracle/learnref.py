"""FP64 CPU oracle of the learned proposal's convolutional encoder.

TEST INFRASTRUCTURE ONLY: tests/ may import this module; the product package
never does.  The encoder has no counterpart in the reference (ImageQuilting.jl
scans every candidate pattern, ``iqsim.jl:206`` via ``utils.jl:5-13``); this
module restates, in numpy float64, the network and the ranking loss that the
HIP training kernels implement in float32, so that their forward values and
gradients can be checked.  Its own analytic gradient is pinned by central
differences (tests/test_oracle.py).

Network (all convolutions 3x3x3, zero outside the training image):
    h1 = relu(conv(x,  w1) + b1)        1 -> 32 channels
    h2 = relu(conv(h1, w2) + b2)       32 -> 32
    f  = x + conv(h2, w3) + b3         32 -> 1   (residual)
    conv(a, w)[o, p] = sum_i sum_k w[o, i, k] a[i, p + d_k],
    tap k = (dx + 1) + 3 (dy + 1) + 9 (dz + 1), d_k = (dx, dy, dz) in {-1, 0, 1}^3.
Embeddings of one overlap-mask shape (V voxels at tile offsets toff, W: 64 x V):
    candidate c:  e_c = W f[c + toff]       query q (V values):  e_q = W q
Approximate distance (the bank scan's):  dd_c = A2_c - 2 e_c . e_q  (+ B2,
constant per query).  Loss of a query over its candidate set S with exact
distances D and temperature tau:
    KL(p* || p~),  p*_c ~ exp(-(D_c - min D) / tau),  p~_c ~ exp(-(dd_c - min dd) / tau)
averaged over the queries.

Parameter vector (float, this order; the HIP library uses the same packing):
    w1[32][27], b1[32], w2[32][32][27], b2[32], w3[32][27], b3[1], W[64][V]
"""
from __future__ import annotations

import numpy as np

C = 32          # channels of the hidden layers
EMB = 64        # embedding width
NCONV = C * 27 + C + C * C * 27 + C + C * 27 + 1  # 29,441 convolution parameters

TAPS = np.array([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])


def nparams(V: int) -> int:
    return NCONV + EMB * V


def unpack(theta: np.ndarray, V: int):
    t = np.asarray(theta, dtype=np.float64)
    assert t.size == nparams(V)
    o = 0

    def take(n, shape):
        nonlocal o
        a = t[o:o + n].reshape(shape)
        o += n
        return a

    w1 = take(C * 27, (C, 1, 27))
    b1 = take(C, (C,))
    w2 = take(C * C * 27, (C, C, 27))
    b2 = take(C, (C,))
    w3 = take(C * 27, (1, C, 27))
    b3 = take(1, (1,))
    W = take(EMB * V, (EMB, V))
    return w1, b1, w2, b2, w3, b3, W


def pack(w1, b1, w2, b2, w3, b3, W) -> np.ndarray:
    return np.concatenate([np.ravel(w1), np.ravel(b1), np.ravel(w2), np.ravel(b2), np.ravel(w3), np.ravel(b3),
                           np.ravel(W)])


def _shift(a: np.ndarray, d) -> np.ndarray:
    """s[..., p] = a[..., p + d], zero where p + d leaves the grid (last three axes)."""
    out = np.zeros_like(a)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    for ax, dd in zip(range(a.ndim - 3, a.ndim), d):
        n = a.shape[ax]
        if dd >= 0:
            src[ax], dst[ax] = slice(dd, n), slice(0, n - dd)
        else:
            src[ax], dst[ax] = slice(0, n + dd), slice(-dd, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def conv(a: np.ndarray, w: np.ndarray) -> np.ndarray:
    """a: (Ci, X, Y, Z), w: (Co, Ci, 27) -> (Co, X, Y, Z)."""
    out = np.zeros((w.shape[0],) + a.shape[1:])
    for k, d in enumerate(TAPS):
        out += np.einsum("oi,ixyz->oxyz", w[:, :, k], _shift(a, d))
    return out


def conv_adjoint_input(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of sum(g * conv(a, w)) with respect to a: (Co,...) -> (Ci,...)."""
    out = np.zeros((w.shape[1],) + g.shape[1:])
    for k, d in enumerate(TAPS):
        out += np.einsum("oi,oxyz->ixyz", w[:, :, k], _shift(g, tuple(-np.asarray(d))))
    return out


def conv_adjoint_weight(g: np.ndarray, a: np.ndarray) -> np.ndarray:
    """Gradient of sum(g * conv(a, w)) with respect to w: (Co, Ci, 27)."""
    out = np.zeros((g.shape[0], a.shape[0], 27))
    for k, d in enumerate(TAPS):
        out[:, :, k] = np.einsum("oxyz,ixyz->oi", g, _shift(a, d))
    return out


def forward(x: np.ndarray, theta: np.ndarray, V: int):
    """x: training image (X, Y, Z) -> (h1, h2, f)."""
    w1, b1, w2, b2, w3, b3, _ = unpack(theta, V)
    x4 = np.asarray(x, dtype=np.float64)[None]
    z1 = conv(x4, w1) + b1[:, None, None, None]
    h1 = np.maximum(z1, 0.0)
    z2 = conv(h1, w2) + b2[:, None, None, None]
    h2 = np.maximum(z2, 0.0)
    f = x4[0] + conv(h2, w3)[0] + b3[0]
    return h1, h2, f


def windows(f: np.ndarray, cands: np.ndarray, toff: np.ndarray) -> np.ndarray:
    """F[n, v] = f[cand n + toff v]; cands: (n, 3) corners, toff: (V, 3) tile offsets."""
    c = np.asarray(cands)[:, None, :] + np.asarray(toff)[None, :, :]
    return f[c[..., 0], c[..., 1], c[..., 2]]


def loss_and_grad(x, theta, toff, A2, Q, cands, D, tau):
    """Ranking loss and its gradient with respect to theta.

    x: training image (X, Y, Z); toff: (V, 3) mask voxels as tile offsets;
    A2: dict-like or callable giving A2 of a candidate corner, here an array
    (nq, ns) of each listed candidate's A2; Q: (nq, V) query values;
    cands: (nq, ns, 3) candidate corners; D: (nq, ns) exact distances;
    tau: (nq,) temperatures.  Returns (loss, grad, detail)."""
    V = len(toff)
    w1, b1, w2, b2, w3, b3, W = unpack(theta, V)
    h1, h2, f = forward(x, theta, V)
    nq, ns = np.asarray(D).shape
    Q = np.asarray(Q, dtype=np.float64)
    eq = Q @ W.T                                        # (nq, 64)
    loss = 0.0
    gW = np.zeros_like(W)
    gf = np.zeros_like(f)
    for i in range(nq):
        F = windows(f, cands[i], toff)                  # (ns, V)
        ec = F @ W.T                                    # (ns, 64)
        dd = np.asarray(A2[i], dtype=np.float64) - 2.0 * ec @ eq[i]
        t = float(tau[i])
        ls = -(dd - dd.min()) / t
        lp = ls - np.log(np.exp(ls).sum())
        lt = -(np.asarray(D[i], dtype=np.float64) - np.min(D[i])) / t
        lpt = lt - np.log(np.exp(lt).sum())
        pt = np.exp(lpt)
        loss += float((pt * (lpt - lp)).sum()) / nq
        gdd = (pt - np.exp(lp)) / t / nq                # dL / d dd_c
        gec = -2.0 * gdd[:, None] * eq[i][None, :]      # dL / d e_c
        geq = -2.0 * (gdd[:, None] * ec).sum(axis=0)    # dL / d e_q
        gW += gec.T @ F + np.outer(geq, Q[i])
        gF = gec @ W                                    # dL / d F  (ns, V)
        c = np.asarray(cands[i])[:, None, :] + np.asarray(toff)[None, :, :]
        np.add.at(gf, (c[..., 0], c[..., 1], c[..., 2]), gF)
    # f = x + conv(h2, w3) + b3
    g3 = gf[None]
    gw3 = conv_adjoint_weight(g3, h2)
    gb3 = np.array([g3.sum()])
    gh2 = conv_adjoint_input(g3, w3) * (h2 > 0)
    gw2 = conv_adjoint_weight(gh2, h1)
    gb2 = gh2.sum(axis=(1, 2, 3))
    gh1 = conv_adjoint_input(gh2, w2) * (h1 > 0)
    gw1 = conv_adjoint_weight(gh1, np.asarray(x, dtype=np.float64)[None])
    gb1 = gh1.sum(axis=(1, 2, 3))
    grad = pack(gw1, gb1, gw2, gb2, gw3, gb3, gW)
    return loss, grad, dict(f=f, eq=eq)


def loss_only(x, theta, toff, A2, Q, cands, D, tau) -> float:
    V = len(toff)
    W = unpack(theta, V)[6]
    _, _, f = forward(x, theta, V)
    eq = np.asarray(Q, dtype=np.float64) @ W.T
    loss = 0.0
    nq = len(tau)
    for i in range(nq):
        ec = windows(f, cands[i], toff) @ W.T
        dd = np.asarray(A2[i], dtype=np.float64) - 2.0 * ec @ eq[i]
        t = float(tau[i])
        ls = -(dd - dd.min()) / t
        lp = ls - np.log(np.exp(ls).sum())
        lt = -(np.asarray(D[i], dtype=np.float64) - np.min(D[i])) / t
        lpt = lt - np.log(np.exp(lt).sum())
        loss += float((np.exp(lpt) * (lpt - lp)).sum()) / nq
    return loss
