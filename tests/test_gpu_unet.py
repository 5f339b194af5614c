"""The learned-preconditioner network on the GPU (BF16 MFMA forward and
backward, Adam) against the CPU oracle (oracle/unet_oracle.py), and the
flexible-CG solve with the network attached.

The device keeps conv weights, activations and activation gradients in BF16
and accumulates in FP32 (biases and the output layer stay FP32).  The oracle
restates that with a BF16 rounding at the same storage points (rnd=).  The
forward output must match it to relative L2 1e-2.  The gradients cannot be
held as tightly: a different FP32 summation order flips an occasional BF16
rounding (or ReLU mask), and the backward pass amplifies those flips.  In
the oracle itself, perturbing every convolution sum by 1e-6 relative (FP32
round-off) moves the weight gradients by up to 1.6e-2 (four and five
levels) and up to 1.6e-1 at the bottom levels of the six-level network, so
the bounds per depth are 3e-2 / 5e-2 / 3e-1 (relative L2) with cosine
0.999 / 0.998 / 0.95.  Against the exact (unrounded) oracle the BF16
storage itself costs 3-9 % per layer, checked as a sanity bound (cosine
0.98, 0.95 for six levels)."""
import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.ims import ImsLinearSettings
from modflow6_amd.learned import UNet, features, init_params, layer_shapes
from oracle import oracle as O
from oracle import unet_oracle as U
from tests.helpers import device_solver, oracle_system

pytestmark = pytest.mark.gpu


def to_oracle(flat, nin, ch, wround=None):
    """Flat parameters -> oracle (W, b) pairs; wround rounds the conv
    weights only (the device packs those to BF16)."""
    wr = wround or (lambda a: a)
    out, o = [], 0
    for a, b in layer_shapes(nin, ch):
        w = wr(flat[o:o + 27 * a * b]).reshape(27, a, b).astype(np.float64); o += 27 * a * b
        bb = flat[o:o + b].astype(np.float64); o += b
        out.append((w, bb))
    wo = flat[o:o + ch[0]].reshape(ch[0], 1).astype(np.float64); o += ch[0]
    out.append((wo, flat[o:o + 1].astype(np.float64)))
    return out


def from_oracle(grads):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in grads])


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def bf16d(a):
    """BF16 rounding (via FP32, nearest-even) of a float64 array, as float64."""
    return bf16(a).astype(np.float64)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def cos(a, b):
    return float(np.dot(a.ravel(), b.ravel()) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))


# four levels on even extents; five and six levels with odd extents on the
# way down (13 -> 7 -> 4 -> 2 -> 1 rows), so the partial pooling windows and
# the cropped upsampling are covered
SHAPES = [((2, 16, 16), (16, 16, 32, 32)), ((3, 16, 24), (16, 32, 64, 64)),
          ((2, 13, 21), (16, 32, 32, 48, 64)), ((2, 40, 36), (16, 16, 32, 32, 64, 64))]


def sample(shape, seed):
    rng = np.random.default_rng(seed)
    D, H, W = shape
    x = rng.standard_normal((3, D * H * W)).astype(np.float32)
    t = rng.standard_normal(D * H * W).astype(np.float32)
    return x, t


@pytest.mark.parametrize("shape,ch", SHAPES)
def test_unet_forward_matches_oracle(shape, ch):
    D, H, W = shape
    p = init_params(3, ch, np.random.default_rng(3))
    p[::7] += 0.05          # non-zero biases too
    net = UNet(D, H, W, 3, ch, params=p)
    x, _ = sample(shape, 4)
    y = net.forward(x)
    net.destroy()
    xo = bf16(x).astype(np.float64).T.reshape(D, H, W, 3)
    yr = U.forward(to_oracle(p, 3, ch, bf16d), xo, rnd=bf16d).reshape(-1)
    ye = U.forward(to_oracle(p, 3, ch), xo).reshape(-1)
    print("forward rel (BF16 oracle)", rel(y, yr), "cos", cos(y, yr), "| exact", rel(y, ye))
    assert rel(y, yr) < 1e-2 and cos(y, yr) > 0.9999
    assert cos(y, ye) > 0.98


@pytest.mark.parametrize("shape,ch", SHAPES)
def test_unet_gradients_match_oracle(shape, ch):
    D, H, W = shape
    p = init_params(3, ch, np.random.default_rng(5))
    net = UNet(D, H, W, 3, ch, params=p)
    x, t = sample(shape, 6)
    loss, g = net.grad(x, t)
    net.destroy()
    xo = bf16(x).astype(np.float64).T.reshape(D, H, W, 3)
    to = bf16(t).astype(np.float64).reshape(D, H, W, 1)
    lr_, gr = U.loss_and_grads(to_oracle(p, 3, ch, bf16d), xo, to, rnd=bf16d)
    le, ge = U.loss_and_grads(to_oracle(p, 3, ch), xo, to)
    assert abs(loss - lr_) <= 1e-3 * lr_
    rtol, ctol, etol = {4: (3e-2, 0.999, 0.98), 5: (5e-2, 0.998, 0.98),
                        6: (3e-1, 0.95, 0.95)}[len(ch)]
    o = 0
    for k, ((w, b), (we, be)) in enumerate(zip(gr, ge)):
        nw, nb = w.size, b.size
        gw, gb = g[o:o + nw], g[o + nw:o + nw + nb]
        o += nw + nb
        print(f"layer {k}: weight rel {rel(gw, w.ravel()):.3e} bias rel {rel(gb, b.ravel()):.3e}"
              f" | exact: weight rel {rel(gw, we.ravel()):.3e} cos {cos(gw, we.ravel()):.5f}")
        assert rel(gw, w.ravel()) < rtol and cos(gw, w.ravel()) > ctol
        assert rel(gb, b.ravel()) < rtol
        assert cos(gw, we.ravel()) > etol
    assert o == g.size


def test_unet_training_reduces_loss():
    shape, ch = (2, 16, 16), (16, 16, 32, 32)
    D, H, W = shape
    rng = np.random.default_rng(7)
    # targets a smooth function of the inputs: a learnable map
    n = 8
    data = np.zeros((n, 4, D * H * W), dtype=np.float32)
    for i in range(n):
        x = rng.standard_normal((3, D, H, W))
        data[i, :3] = x.reshape(3, -1)
        tt = 0.5 * x[0] + np.roll(x[1], 1, axis=2) - 0.3 * x[2]
        data[i, 3] = tt.reshape(-1)
    net = UNet(D, H, W, 3, ch, seed=1)
    net.set_data(data)
    loss = net.train(nsteps=120, batch=2, lr=2e-3, seed=0)
    net.destroy()
    print("loss first/last", loss[:3], loss[-3:])
    assert np.all(np.isfinite(loss))
    assert loss[-10:].mean() < 0.5 * loss[:10].mean()


def test_fcg_with_untrained_network_converges():
    """A random network must not break the solve: the energy-optimal step
    and the fallback keep flexible CG converging to the ILU0-PCG heads."""
    case = configs.steady_small(nlay=2, nrow=32, ncol=32, nwel=5)
    sim, amat, rhs, x0 = oracle_system(case)
    ref = device_solver(case)
    xr = x0.copy()
    ic, it_ref = ref.solve(1, amat, rhs, xr)
    ref.destroy()
    sv = device_solver(case, preconditioner=1)
    net = UNet(case.nlay, case.nrow, case.ncol, 3, (16, 16, 32, 32), seed=2)
    sv.attach_unet(net, features(case))
    xl = x0.copy()
    icl, it_l = sv.solve(1, amat, rhs, xl)
    nl = sv.learned_iterations()
    sv.destroy()
    net.destroy()
    print("ILU0-PCG", it_ref, "FCG+net", it_l, "learned iterations", nl)
    assert ic == 1 and icl == 1
    assert np.max(np.abs(xl - xr)) < 10 * case.ims.inner_dvclose


def test_fcg_without_network_matches_pcg():
    case = configs.steady_small(nlay=2, nrow=32, ncol=32, nwel=5)
    sim, amat, rhs, x0 = oracle_system(case)
    ref = device_solver(case)
    xr = x0.copy()
    ic, it_ref = ref.solve(1, amat, rhs, xr)
    ref.destroy()
    sv = device_solver(case, preconditioner=1)
    xl = x0.copy()
    icl, it_l = sv.solve(1, amat, rhs, xl)
    sv.destroy()
    print("ILU0-PCG", it_ref, "ILU0-FCG", it_l)
    assert icl == 1 and abs(it_l - it_ref) <= max(3, it_ref // 20)
    assert np.max(np.abs(xl - xr)) < 10 * case.ims.inner_dvclose


def test_trained_network_gain_over_flexible_cg_without_it():
    """What the trained network gains: a 6-level network trained on the
    device on ILU0-PCG (residual, error) pairs of eight seeded 4 x 64 x 64
    fields, applied to three unseen fields inside flexible CG, against the
    same flexible CG without it.  The gain (ILU0-only iterations / learned
    iterations) is printed; at a tight tolerance it is below one (a
    correction must be accurate to ~0.1 % per cell to pay, a BF16 network
    is not).  The test holds what the method guarantees: training lowers the
    loss and every solve, with or without the network, converges to the
    ILU0-PCG heads (the FP64 true residual decides)."""
    from modflow6_amd import learned as LN
    shape = (4, 64, 64)
    data = LN.make_dataset(list(range(300, 308)), shape, log=lambda *a: None)
    net = LN.UNet(*shape, nin=3, channels=LN.CHANNELS, seed=0)
    net.set_data(data)
    losses = net.train(240, 4, 1e-3, seed=0)
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-20:]) < np.mean(losses[:5])
    gains = []
    for seed in (900, 901, 902):
        r = LN.solve_pair(LN.learned_case(seed, *shape), net, fcg_m=8)
        assert r["pcg"]["converged"] and r["fcg_ilu0"]["converged"]
        assert r["fcg_learned"]["converged"]
        assert r["fcg_learned"]["max_head_diff_vs_pcg_m"] < 1e-5
        base, lrn = r["fcg_ilu0"]["iterations"], r["fcg_learned"]["iterations"]
        gains.append(base / lrn)
        print(f"seed {seed}: PCG {r['pcg']['iterations']}, FCG+ILU0 {base}, FCG+net {lrn} "
              f"({r['fcg_learned']['learned_iterations']} learned), gain {base / lrn:.2f}")
    net.destroy()
    print("iteration gain of the trained network:", [f"{g:.2f}" for g in gains])
