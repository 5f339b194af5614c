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
    def conv(h, w, b):
        return F.conv3d(h, w.reshape(3, 3, 3, w.shape[1], w.shape[2]).permute(4, 3, 0, 1, 2),
                        b, padding=1)
    h = xt
    acts = []
    for lvl in range(4):
        if lvl > 0:
            h = F.avg_pool3d(h, (1, 2, 2))
        h = F.relu(conv(h, *tp[2 * lvl]))
        h = F.relu(conv(h, *tp[2 * lvl + 1]))
        acts.append(h)
    u = acts[3]
    for lvl, idx in ((2, 8), (1, 9), (0, 10)):
        cat = torch.cat([F.interpolate(u, scale_factor=(1, 2, 2), mode="nearest"), acts[lvl]], 1)
        u = F.relu(conv(cat, *tp[idx]))
    wo, bo = tp[11]
    return torch.einsum("ncdhw,co->nodhw", u, wo) + bo.reshape(1, -1, 1, 1, 1)


@pytest.mark.parametrize("shape,ch", [((2, 8, 8), (4, 6, 8, 8)), ((3, 16, 8), (16, 16, 32, 32))])
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
