"""CPU oracle of the learned preconditioner network (test infrastructure only:
imported by tests/ to check the HIP forward and backward kernels).

The network has no counterpart in the reference; this file states its
mathematics in plain numpy, independently of the HIP implementation:

* tensors are channels-last, shape (D, H, W, C) = (nlay, nrow, ncol, C);
* conv3(x, W, b): 3x3x3 cross-correlation with zero padding 1,
  y[d,h,w,o] = b[o] + sum_{t=(i,j,k)} sum_c x[d+i-1, h+j-1, w+k-1, c] W[t,c,o]
  with tap index t = 9 i + 3 j + k;
* pool: mean over 1x2x2 blocks (H, W -> ceil(H / 2), ceil(W / 2); a block
  on the last row / column of an odd extent averages the 1 or 2 cells it
  has); up: nearest 1x2x2 repeat cropped to the fine extent;
* levels 0..L-1 with channels ch[0..L-1]:
    a0 = relu(conv(relu(conv(in))))            encoder, level 0
    a_l = relu(conv(relu(conv(pool(a_{l-1})))))  levels 1..L-1
    u_{L-2} = relu(conv(cat(up(a_{L-1}), a_{L-2})))   decoder (cat: up first)
    u_l = relu(conv(cat(up(u_{l+1}), a_l)))    l = L-3..0
    y  = u0 . w_out + b_out                    1x1x1 output conv
* loss = sum (y - t)^2 / sum t^2 (relative squared error of the batch).

Parameters are kept as a list of (W, b) pairs in the order of LAYERS.
"""
from __future__ import annotations

import numpy as np

# (name, cin, cout) of the conv layers in parameter order
def layer_shapes(cin, ch):
    L = len(ch)
    out = []
    for l in range(L):
        out += [(f"e{l}a", cin if l == 0 else ch[l - 1], ch[l]), (f"e{l}b", ch[l], ch[l])]
    out += [(f"d{l}", ch[l + 1] + ch[l], ch[l]) for l in range(L - 2, -1, -1)]
    return out


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


def _blocks(H, W):
    """(a, b, rows, cols): the fine cells (2h + a, 2w + b) of each coarse
    cell that exist, as slices of the fine and coarse grids."""
    Hc, Wc = (H + 1) // 2, (W + 1) // 2
    for a in (0, 1):
        for b in (0, 1):
            nh, nw = len(range(a, H, 2)), len(range(b, W, 2))
            yield (slice(a, None, 2), slice(b, None, 2)), (slice(0, nh), slice(0, nw)), (Hc, Wc)


def _count(H, W):
    Hc, Wc = (H + 1) // 2, (W + 1) // 2
    n = np.zeros((Hc, Wc))
    for _, (rc, cc), _ in _blocks(H, W):
        n[rc, cc] += 1.0
    return n[None, :, :, None]


def pool(x):
    D, H, W, C = x.shape
    s = np.zeros((D, (H + 1) // 2, (W + 1) // 2, C))
    for (rf, cf), (rc, cc), _ in _blocks(H, W):
        s[:, rc, cc] += x[:, rf, cf]
    return s / _count(H, W)


def pool_bwd(dy, H, W):
    g = dy / _count(H, W)
    dx = np.zeros((dy.shape[0], H, W, dy.shape[-1]))
    for (rf, cf), (rc, cc), _ in _blocks(H, W):
        dx[:, rf, cf] = g[:, rc, cc]
    return dx


def up(x, H, W):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :H, :W]


def up_bwd(dy):
    D, H, W, C = dy.shape
    s = np.zeros((D, (H + 1) // 2, (W + 1) // 2, C))
    for (rf, cf), (rc, cc), _ in _blocks(H, W):
        s[:, rc, cc] += dy[:, rf, cf]
    return s


def relu(x):
    return np.maximum(x, 0.0)


def _named(params, cin):
    nlev = len(params) // 3
    ch = [p[0].shape[2] for p in params[0:2 * nlev:2]]
    names = [n for n, _, _ in layer_shapes(cin, ch)]
    return ch, names, dict(zip(names, params[:-1]))


def _identity(a):
    return a


def forward(params, x, keep=False, rnd=None):
    """Output y (D, H, W, 1).  rnd, when given, is applied to every stored
    activation (pooled inputs, conv outputs): with a BF16 rounding it
    restates a device that keeps activations in BF16."""
    R = rnd or _identity
    ch, _, P = _named(params, x.shape[-1])
    nlev = len(ch)
    st = {}
    h = x
    acts = []
    for lvl in range(nlev):
        if lvl > 0:
            h = R(pool(h))
        st[f"in{lvl}"] = h
        za = conv3(h, *P[f"e{lvl}a"]); ha = R(relu(za))
        zb = conv3(ha, *P[f"e{lvl}b"]); hb = R(relu(zb))
        st[f"h{lvl}a"], st[f"h{lvl}b"] = ha, hb
        acts.append(hb)
        h = hb
    u = acts[-1]
    for lvl in range(nlev - 2, -1, -1):
        a = acts[lvl]
        cat = np.concatenate([up(u, a.shape[1], a.shape[2]), a], axis=-1)
        st[f"cat{lvl}"] = cat
        u = R(relu(conv3(cat, *P[f"d{lvl}"])))
        st[f"u{lvl}"] = u
    wo, bo = params[-1]
    y = u @ wo + bo
    return (y, st) if keep else y


def loss_and_grads(params, x, t, rnd=None):
    """Relative squared error of one sample and its parameter gradients.
    rnd, when given, is applied to the activations (as in forward) and to
    every stored activation gradient (the gradients of the decoder outputs,
    of the concatenations, of both convolutions of each encoder level and of
    the pooled inputs)."""
    R = rnd or _identity
    ch, names, P = _named(params, x.shape[-1])
    y, st = forward(params, x, keep=True, rnd=rnd)
    den = np.sum(t * t)
    loss = np.sum((y - t) ** 2) / den
    dy = 2.0 * (y - t) / den
    G = {}
    wo, bo = params[-1]
    u0 = st["u0"]
    g_out = (u0.reshape(-1, u0.shape[-1]).T @ dy.reshape(-1, 1), dy.sum(axis=(0, 1, 2)))
    du = dy @ wo.T
    nlev = len(ch)
    dacts = {lvl: 0.0 for lvl in range(nlev)}
    for lvl in range(nlev - 1):
        du = R(du * (st[f"u{lvl}"] > 0))
        dcat, dw, db = conv3_bwd(st[f"cat{lvl}"], P[f"d{lvl}"][0], du)
        dcat = R(dcat)
        G[f"d{lvl}"] = (dw, db)
        cu = dcat.shape[-1] - ch[lvl]
        dacts[lvl] = dacts[lvl] + dcat[..., cu:]
        du = up_bwd(dcat[..., :cu])
    dacts[nlev - 1] = dacts[nlev - 1] + du
    dh = None
    for lvl in range(nlev - 1, -1, -1):
        g = dacts[lvl] if dh is None else dacts[lvl] + dh
        g = R(g * (st[f"h{lvl}b"] > 0))
        dha, dw, db = conv3_bwd(st[f"h{lvl}a"], P[f"e{lvl}b"][0], g)
        G[f"e{lvl}b"] = (dw, db)
        dha = R(dha * (st[f"h{lvl}a"] > 0))
        din, dw, db = conv3_bwd(st[f"in{lvl}"], P[f"e{lvl}a"][0], dha)
        G[f"e{lvl}a"] = (dw, db)
        if lvl > 0:
            hp = st[f"h{lvl - 1}b"]
            dh = pool_bwd(R(din), hp.shape[1], hp.shape[2])
    grads = [G[n] for n in names] + [g_out]
    return loss, grads
