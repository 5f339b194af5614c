"""CPU oracle of the learned preconditioner network (test infrastructure only:
imported by tests/ to check the HIP forward and backward kernels).

The network has no counterpart in the reference; this file states its
mathematics in plain numpy, independently of the HIP implementation:

* tensors are channels-last, shape (D, H, W, C) = (nlay, nrow, ncol, C);
* conv3(x, W, b): 3x3x3 cross-correlation with zero padding 1,
  y[d,h,w,o] = b[o] + sum_{t=(i,j,k)} sum_c x[d+i-1, h+j-1, w+k-1, c] W[t,c,o]
  with tap index t = 9 i + 3 j + k;
* pool: mean over 1x2x2 blocks (H, W halved); up: nearest 1x2x2 repeat;
* levels 0..3 with channels ch[0..3]:
    a0 = relu(conv(relu(conv(in))))            encoder, level 0
    a1 = relu(conv(relu(conv(pool(a0)))))      level 1, likewise a2, a3
    u2 = relu(conv(cat(up(a3), a2)))           decoder (cat: up channels first)
    u1 = relu(conv(cat(up(u2), a1)))
    u0 = relu(conv(cat(up(u1), a0)))
    y  = u0 . w_out + b_out                    1x1x1 output conv
* loss = sum (y - t)^2 / sum t^2 (relative squared error of the batch).

Parameters are kept as a list of (W, b) pairs in the order of LAYERS.
"""
from __future__ import annotations

import numpy as np

# (name, cin_source, cout_index); conv layers in evaluation order
def layer_shapes(cin, ch):
    c0, c1, c2, c3 = ch
    return [("e0a", cin, c0), ("e0b", c0, c0),
            ("e1a", c0, c1), ("e1b", c1, c1),
            ("e2a", c1, c2), ("e2b", c2, c2),
            ("e3a", c2, c3), ("e3b", c3, c3),
            ("d2", c3 + c2, c2), ("d1", c2 + c1, c1), ("d0", c1 + c0, c0)]


def init_params(cin, ch, rng):
    ps = []
    for _, a, b in layer_shapes(cin, ch):
        w = rng.standard_normal((27, a, b)) * np.sqrt(2.0 / (27 * a))
        ps.append((w, np.zeros(b)))
    ps.append((rng.standard_normal((ch[0], 1)) * np.sqrt(1.0 / ch[0]), np.zeros(1)))
    return ps


def conv3(x, w, b):
    D, H, W, C = x.shape
    xp = np.zeros((D + 2, H + 2, W + 2, C))
    xp[1:-1, 1:-1, 1:-1] = x
    y = np.zeros((D, H, W, w.shape[2])) + b
    for t in range(27):
        i, j, k = t // 9, (t // 3) % 3, t % 3
        y += xp[i:i + D, j:j + H, k:k + W] @ w[t]
    return y


def conv3_bwd(x, w, dy):
    """Gradients of conv3 with respect to x, w and b."""
    D, H, W, C = x.shape
    xp = np.zeros((D + 2, H + 2, W + 2, C))
    xp[1:-1, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for t in range(27):
        i, j, k = t // 9, (t // 3) % 3, t % 3
        sl = xp[i:i + D, j:j + H, k:k + W]
        dw[t] = sl.reshape(-1, C).T @ dy.reshape(-1, dy.shape[-1])
        dxp[i:i + D, j:j + H, k:k + W] += dy @ w[t].T
    return dxp[1:-1, 1:-1, 1:-1], dw, dy.reshape(-1, dy.shape[-1]).sum(0)


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(dy):
    D, H, W, C = dy.shape
    return dy.reshape(D, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def relu(x):
    return np.maximum(x, 0.0)


def _named(params, cin):
    ch = [p[0].shape[2] for p in params[0:8:2]]
    names = [n for n, _, _ in layer_shapes(cin, ch)]
    return ch, names, dict(zip(names, params[:-1]))


def forward(params, x, keep=False):
    _, _, P = _named(params, x.shape[-1])
    st = {}
    h = x
    acts = []
    for lvl in range(4):
        if lvl > 0:
            h = pool(h)
        st[f"in{lvl}"] = h
        za = conv3(h, *P[f"e{lvl}a"]); ha = relu(za)
        zb = conv3(ha, *P[f"e{lvl}b"]); hb = relu(zb)
        st[f"h{lvl}a"], st[f"h{lvl}b"] = ha, hb
        acts.append(hb)
        h = hb
    u = acts[3]
    for lvl in (2, 1, 0):
        cat = np.concatenate([up(u), acts[lvl]], axis=-1)
        st[f"cat{lvl}"] = cat
        u = relu(conv3(cat, *P[f"d{lvl}"]))
        st[f"u{lvl}"] = u
    wo, bo = params[-1]
    y = u @ wo + bo
    return (y, st) if keep else y


def loss_and_grads(params, x, t):
    """Relative squared error of one sample and its parameter gradients."""
    ch, names, P = _named(params, x.shape[-1])
    y, st = forward(params, x, keep=True)
    den = np.sum(t * t)
    loss = np.sum((y - t) ** 2) / den
    dy = 2.0 * (y - t) / den
    G = {}
    wo, bo = params[-1]
    u0 = st["u0"]
    g_out = (u0.reshape(-1, u0.shape[-1]).T @ dy.reshape(-1, 1), dy.sum(axis=(0, 1, 2)))
    du = dy @ wo.T
    dacts = {lvl: 0.0 for lvl in range(4)}
    for lvl in (0, 1, 2):
        du = du * (st[f"u{lvl}"] > 0)
        dcat, dw, db = conv3_bwd(st[f"cat{lvl}"], P[f"d{lvl}"][0], du)
        G[f"d{lvl}"] = (dw, db)
        cu = dcat.shape[-1] - ch[lvl]
        dacts[lvl] = dacts[lvl] + dcat[..., cu:]
        du = up_bwd(dcat[..., :cu])
    dacts[3] = dacts[3] + du
    dh = None
    for lvl in (3, 2, 1, 0):
        g = dacts[lvl] if dh is None else dacts[lvl] + dh
        g = g * (st[f"h{lvl}b"] > 0)
        dha, dw, db = conv3_bwd(st[f"h{lvl}a"], P[f"e{lvl}b"][0], g)
        G[f"e{lvl}b"] = (dw, db)
        dha = dha * (st[f"h{lvl}a"] > 0)
        din, dw, db = conv3_bwd(st[f"in{lvl}"], P[f"e{lvl}a"][0], dha)
        G[f"e{lvl}a"] = (dw, db)
        dh = pool_bwd(din) if lvl > 0 else din
    grads = [G[n] for n in names] + [g_out]
    return loss, grads
