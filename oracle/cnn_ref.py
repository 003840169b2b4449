"""numpy reference of the learned-window CNN — TEST INFRASTRUCTURE ONLY.

Checker for the HIP kernels in libbicos_amd/csrc/cnn.hpp (forward pass,
correlation, loss and every gradient); imported by tests/ only. The network
has no counterpart in the reference library: this file restates the
network's definition (written out in the header of cnn.hpp) in float64 /
float32 numpy, and its backward pass is itself pinned by finite differences
(tests/test_cnn_ref.py).
"""
from __future__ import annotations

import numpy as np

C, IN, DS, K, D0 = 32, 4, 4, 80, -8
W1 = C * IN * 9
WL = C * C * 9


def w_off(l):
    return 0 if l == 0 else W1 + C + (l - 1) * (WL + C)


def b_off(l):
    return w_off(l) + (W1 if l == 0 else WL)


PARAMS = w_off(4)


def bf16(x):
    """round-to-nearest-even to bfloat16, returned as float32"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for l in range(4):
        cin = IN if l == 0 else C
        n = (W1 if l == 0 else WL)
        p[w_off(l):w_off(l) + n] = rng.normal(0, np.sqrt(2.0 / (cin * 9)), n)
    return p


def layers(p, bf=True):
    out = []
    for l in range(4):
        cin = IN if l == 0 else C
        w = p[w_off(l):w_off(l) + (W1 if l == 0 else WL)].reshape(C, cin, 3, 3).astype(np.float64)
        if bf:
            w = bf16(w.astype(np.float32)).astype(np.float64)
        b = p[b_off(l):b_off(l) + C].astype(np.float64)
        out.append((w, b))
    return out


def cnn_input(frames, in_scale):
    """frames: (>=4, H, W) -> (4, H//4, W//4): 4x4 box means minus their mean over the 4 frames"""
    f = frames[:IN].astype(np.float64)
    Hq, Wq = f.shape[1] // DS, f.shape[2] // DS
    q = f[:, :Hq * DS, :Wq * DS].reshape(IN, Hq, DS, Wq, DS).mean(axis=(2, 4))
    return (q - q.mean(0, keepdims=True)) * in_scale


def conv(x, w, b):
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    y = np.zeros((w.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("oc,chw->ohw", w[:, :, ky, kx], xp[:, ky:ky + H, kx:kx + W])
    return y + b[:, None, None]


def conv_bwd(x, w, dy):
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = np.einsum("ohw,chw->oc", dy, xp[:, ky:ky + H, kx:kx + W])
            dxp[:, ky:ky + H, kx:kx + W] += np.einsum("oc,ohw->chw", w[:, :, ky, kx], dy)
    return dxp[:, 1:-1, 1:-1], dw, dy.sum(axis=(1, 2))


def features(x0, L, bf=True):
    """activations x0, a1, a2, a3, f; with bf every one is rounded to BF16
    (the forward pass of the HIP kernels); the backward pass treats the
    rounding as the identity"""
    rnd = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if bf else (lambda a: a)
    h = rnd(x0)
    acts = [h]
    for l, (w, b) in enumerate(L):
        z = conv(h, w, b)
        h = rnd(np.maximum(z, 0) if l < 3 else z)
        acts.append(h)
    return acts


def corr(fl, fr, bins=K):
    Cn, Hq, Wq = fl.shape
    out = np.full((bins, Hq, Wq), -1e30)
    for k in range(bins):
        d = D0 + k
        for X in range(Wq):
            xr = X - d
            if 0 <= xr < Wq:
                out[k, :, X] = (fl[:, :, X] * fr[:, :, xr]).sum(0) / np.sqrt(C)
    return out


def softmax_ce(Cv, label):
    """label: (Hq, Wq) full-resolution disparity or NaN -> (loss sum, count, dC unscaled)"""
    Kn, Hq, Wq = Cv.shape
    valid = Cv > -1e29
    mx = np.where(valid, Cv, -np.inf).max(0)
    e = np.where(valid, np.exp(Cv - mx), 0.0)
    lz = mx + np.log(e.sum(0))
    p = np.where(valid, np.exp(Cv - lz), 0.0)
    pos = label / DS - D0
    ok = np.isfinite(label) & (pos >= 0) & (pos <= Kn - 1)
    k0 = np.floor(np.where(ok, pos, 0)).astype(int)
    w1 = np.where(ok, pos - k0, 0.0)
    # a cell whose target bins fall outside the right image is unlabelled
    Xs = np.arange(Wq)[None, :]
    xr0 = Xs - (D0 + k0)
    ok &= (xr0 >= 0) & (xr0 < Wq) & ((w1 == 0) | (xr0 - 1 >= 0))
    tgt = np.zeros_like(Cv)
    Y, X = np.nonzero(ok)
    tgt[k0[Y, X], Y, X] = 1 - w1[Y, X]
    hi = k0[Y, X] + 1
    m = hi < Kn
    tgt[hi[m], Y[m], X[m]] += w1[Y[m], X[m]]
    loss = -(tgt * np.where(valid, Cv - lz, 0.0)).sum()
    dC = np.where(ok[None], p - tgt, 0.0)
    return loss, int(ok.sum()), dC


def corr_bwd(fl, fr, dC):
    Cn, Hq, Wq = fl.shape
    dfl = np.zeros_like(fl)
    dfr = np.zeros_like(fr)
    s = 1 / np.sqrt(C)
    for k in range(dC.shape[0]):
        d = D0 + k
        for X in range(Wq):
            xr = X - d
            if 0 <= xr < Wq:
                g = dC[k, :, X] * s
                dfl[:, :, X] += g * fr[:, :, xr]
                dfr[:, :, xr] += g * fl[:, :, X]
    return dfl, dfr


def loss_and_grad(framesL, framesR, label, p, in_scale, bf=True, bins=K):
    """mean cross-entropy over labelled cells and its gradient w.r.t. the flat
    parameters (weights and activations through their BF16 values when bf)"""
    L = layers(p, bf)
    aL = features(cnn_input(framesL, in_scale), L, bf)
    aR = features(cnn_input(framesR, in_scale), L, bf)
    Cv = corr(aL[-1], aR[-1], bins)
    loss, cnt, dC = softmax_ce(Cv, label)
    cnt = max(cnt, 1)
    dC /= cnt
    g = np.zeros(PARAMS)
    dL, dR = corr_bwd(aL[-1], aR[-1], dC)
    for acts, d in ((aL, dL), (aR, dR)):
        for l in range(3, -1, -1):
            if l < 3:
                d = d * (acts[l + 1] > 0)
            dx, dw, db = conv_bwd(acts[l], L[l][0], d)
            g[w_off(l):w_off(l) + dw.size] += dw.ravel()
            g[b_off(l):b_off(l) + C] += db
            d = dx
    return loss / cnt, g, Cv
