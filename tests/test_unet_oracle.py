"""Pins the CPU oracle of the learned-preconditioner network
(oracle/unet_oracle.py) against torch's conv3d / avg_pool3d / interpolate
and autograd on the CPU: forward output and every parameter gradient agree
to round-off.  (The network has no counterpart in the reference.)"""
import numpy as np
import pytest

from oracle import unet_oracle as U

torch = pytest.importorskip("torch")
F = torch.nn.functional


def torch_net(tp, xt):
    nlev = len(tp) // 3

    def conv(h, w, b):
        return F.conv3d(h, w.reshape(3, 3, 3, w.shape[1], w.shape[2]).permute(4, 3, 0, 1, 2),
                        b, padding=1)
    h = xt
    acts = []
    for lvl in range(nlev):
        if lvl > 0:
            # ceil_mode: a window on the last row / column of an odd extent
            # averages the cells it covers
            h = F.avg_pool3d(h, (1, 2, 2), ceil_mode=True)
        h = F.relu(conv(h, *tp[2 * lvl]))
        h = F.relu(conv(h, *tp[2 * lvl + 1]))
        acts.append(h)
    u = acts[-1]
    for i, lvl in enumerate(range(nlev - 2, -1, -1)):
        a = acts[lvl]
        up = u.repeat_interleave(2, dim=3).repeat_interleave(2, dim=4)
        cat = torch.cat([up[:, :, :, :a.shape[3], :a.shape[4]], a], 1)
        u = F.relu(conv(cat, *tp[2 * nlev + i]))
    wo, bo = tp[-1]
    return torch.einsum("ncdhw,co->nodhw", u, wo) + bo.reshape(1, -1, 1, 1, 1)


@pytest.mark.parametrize("shape,ch", [((2, 8, 8), (4, 6, 8, 8)), ((3, 16, 8), (16, 16, 32, 32)),
                                      ((2, 13, 11), (4, 6, 8, 8, 8)),
                                      ((2, 21, 9), (4, 4, 6, 6, 8)),
                                      ((1, 40, 36), (4, 4, 4, 6, 6, 8))])
def test_unet_oracle_matches_torch_autograd(shape, ch):
    rng = np.random.default_rng(1)
    ps = U.init_params(3, ch, rng)
    ps = [(w, rng.standard_normal(b.shape) * 0.1) for w, b in ps]
    x = rng.standard_normal(shape + (3,))
    t = rng.standard_normal(shape + (1,))
    loss, grads = U.loss_and_grads(ps, x, t)
    tp = [(torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in ps]
    y = torch_net(tp, torch.tensor(x).permute(3, 0, 1, 2)[None])
    tt = torch.tensor(t).permute(3, 0, 1, 2)[None]
    L = ((y - tt) ** 2).sum() / (tt ** 2).sum()
    L.backward()
    assert abs(loss - L.item()) <= 1e-12 * abs(loss)
    yo = U.forward(ps, x)
    assert np.max(np.abs(yo - y.detach()[0].permute(1, 2, 3, 0).numpy())) < 1e-12
    for (gw, gb), (w, b) in zip(grads, tp):
        assert np.max(np.abs(gw - w.grad.numpy())) <= 1e-12 * max(1.0, np.max(np.abs(gw)))
        assert np.max(np.abs(gb - b.grad.numpy())) <= 1e-12 * max(1.0, np.max(np.abs(gb)))


def test_pool_of_odd_extent_averages_the_cells_it_covers():
    """Hand-worked 1 x 3 x 3 case: corner block of 4, edge blocks of 2, the
    last corner of 1; the pooling gradient spreads dy / count back."""
    x = np.arange(9.0).reshape(1, 3, 3, 1)
    p = U.pool(x)[0, :, :, 0]
    assert np.array_equal(p, [[(0 + 1 + 3 + 4) / 4, (2 + 5) / 2], [(6 + 7) / 2, 8.0]])
    dx = U.pool_bwd(np.ones((1, 2, 2, 1)), 3, 3)[0, :, :, 0]
    assert np.array_equal(dx, [[0.25, 0.25, 0.5], [0.25, 0.25, 0.5], [0.5, 0.5, 1.0]])
    assert np.array_equal(U.up_bwd(np.ones((1, 3, 3, 1)))[0, :, :, 0], [[4, 2], [2, 1]])


def test_unet_oracle_gradient_finite_difference():
    """Central differences of the loss along random parameter directions."""
    rng = np.random.default_rng(2)
    ch = (4, 4, 4, 4)
    ps = U.init_params(3, ch, rng)
    ps = [(w, rng.standard_normal(b.shape) * 0.1) for w, b in ps]
    x = rng.standard_normal((2, 8, 8, 3))
    t = rng.standard_normal((2, 8, 8, 1))
    _, grads = U.loss_and_grads(ps, x, t)
    for trial in range(3):
        dirs = [(rng.standard_normal(w.shape), rng.standard_normal(b.shape)) for w, b in ps]
        h = 1e-6
        plus = [(w + h * dw, b + h * db) for (w, b), (dw, db) in zip(ps, dirs)]
        minus = [(w - h * dw, b - h * db) for (w, b), (dw, db) in zip(ps, dirs)]
        fd = (U.loss_and_grads(plus, x, t)[0] - U.loss_and_grads(minus, x, t)[0]) / (2 * h)
        an = sum(np.sum(gw * dw) + np.sum(gb * db) for (gw, gb), (dw, db) in zip(grads, dirs))
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an))
