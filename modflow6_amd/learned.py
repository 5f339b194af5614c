"""Learned preconditioner: training data from the device ILU0-PCG, training
on the device, and the flexible-CG solve with the network attached.

The network (``mf6hip_unet_*``, a 4-level 3D U-Net in BF16 on MFMA) learns
the correction that the ILU0 step leaves: for a PCG iterate x_k with
residual r_k and error e_k = x* - x_k,

    z_k  = ILU0(r_k),   t_k = r_k - A z_k,   target(t_k) = e_k - z_k,

so that ILU0(r) + net(r - A ILU0(r)) approximates A^-1 r.  Inputs are
scaled by s = rms(t): the network sees [t / s, ln K, ln(K33 / K)] (the two
features standardised per model) and predicts (e - z) / s.

Host code only orchestrates: the network, its training and the solve run in
libmf6hip.so (``mf6hip_unet_train``, ``mf6hip_solve`` with
``preconditioner = 1``).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, configs, fields
from .ims import HipLinearSolver, ImsLinearSettings, dis_structure

# six levels: the coarsest grid of a 256 x 256 layer is 8 x 8, so the
# receptive field spans the model
CHANNELS = (16, 32, 64, 64, 64, 64)
# the learned-preconditioner workload's stated network: four levels of
# 16, 32, 64, 64 channels
WORKLOAD_CHANNELS = (16, 32, 64, 64)


def standardise(a):
    a = np.asarray(a, dtype=np.float64)
    sd = a.std()
    return (a - a.mean()) / (sd if sd > 0 else 1.0)


def features(case):
    """Two standardised planes: ln K and ln(K33 / K), DIS node order."""
    f = np.stack([standardise(np.log(case.k11)),
                  standardise(np.log(case.k33 / case.k11))])
    return np.ascontiguousarray(f, dtype=np.float32)


def learned_field(seed, nlay=16, nrow=256, ncol=256):
    """The seeded draws of one learned-workload field: (K, sigma, correlation
    length, K33 / K)."""
    rng = np.random.default_rng(seed)
    sigma = rng.uniform(0.5, 2.5)
    corr = rng.uniform(4.0, 64.0)
    aniso = 10.0 ** rng.uniform(-2.0, 0.0)
    k = fields.correlated_lognormal((nlay, nrow, ncol), np.log(10.0), sigma,
                                    corr, rng).reshape(-1)
    return k, sigma, corr, aniso


def learned_case(seed, nlay=16, nrow=256, ncol=256, field=None):
    """Steady confined GWF model of the learned-preconditioner workload:
    lognormal K with sigma on [0.5, 2.5], exponential correlation length on
    [4, 64] cells, K33 / K on [0.01, 1]; delr = delc = 50 m, 10 m layers;
    CHD 10 / 0 m on the first / last column and RCH 1e-4 m/d.  field: the
    seed's learned_field when already drawn."""
    k, sigma, corr, aniso = field if field is not None else learned_field(seed, nlay, nrow,
                                                                          ncol)
    case = configs.steady_10m(seed=seed, nlay=nlay, nrow=nrow, ncol=ncol, nwel=1, k11=k)
    case.packages = [p for p in case.packages if p.kind != "WEL"]
    case.k11 = k
    case.k33 = aniso * k
    case.name = f"learned_{nlay}x{nrow}x{ncol}_{seed}"
    case.meta = dict(seed=seed, sigma=sigma, corr_len=corr, aniso=aniso)
    return case


def _field_job(args):
    return learned_field(*args)


def learned_cases(seeds, shape, pool=None):
    """learned_case of every seed in seed order, the K fields drawn by `pool`
    (CPU processes, see field_pool) when given."""
    if pool is None:
        for s in seeds:
            yield learned_case(s, *shape)
        return
    for s, f in zip(seeds, pool.imap(_field_job, [(s, *shape) for s in seeds])):
        yield learned_case(s, *shape, field=f)


def field_pool():
    """A pool of spawned CPU processes for learned_field (they never touch
    the GPU).  Create it before the process first uses the GPU: its workers
    start here, so no process is started afterwards."""
    import multiprocessing as mp
    import os
    n = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else (os.cpu_count() or 1))
    return mp.get_context("spawn").Pool(n) if n > 1 else None


class UNet:
    """ctypes handle of one device network."""

    def __init__(self, nlay, nrow, ncol, nin=3, channels=CHANNELS, params=None,
                 seed=0):
        self.L = L = _lib.load()
        self.h = C.c_void_p()
        ch = (C.c_int * len(channels))(*channels)
        _lib.check(L.mf6hip_unet_create(nlay, nrow, ncol, nin, len(channels), ch,
                                        C.byref(self.h)),
                   "mf6hip_unet_create")
        n = C.c_longlong(0)
        _lib.check(L.mf6hip_unet_num_params(self.h, C.byref(n)), "mf6hip_unet_num_params")
        self.nparam = n.value
        self.shape = (nlay, nrow, ncol)
        self.nin = nin
        self.channels = tuple(channels)
        if params is None:
            params = init_params(nin, channels, np.random.default_rng(seed))
        self.set_params(params)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _lib.check(self.L.mf6hip_unet_set_params(self.h, _lib.fptr(p)), "mf6hip_unet_set_params")

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        _lib.check(self.L.mf6hip_unet_get_params(self.h, _lib.fptr(p)), "mf6hip_unet_get_params")
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(int(np.prod(self.shape)), dtype=np.float32)
        _lib.check(self.L.mf6hip_unet_forward(self.h, _lib.fptr(x), _lib.fptr(y)),
                   "mf6hip_unet_forward")
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float(0.0)
        _lib.check(self.L.mf6hip_unet_grad(self.h, _lib.fptr(x), _lib.fptr(t), C.byref(loss),
                                           _lib.fptr(g)), "mf6hip_unet_grad")
        return loss.value, g

    def set_data(self, samples):
        samples = np.ascontiguousarray(samples, dtype=np.float32)
        n = samples.shape[0]
        _lib.check(self.L.mf6hip_unet_set_data(self.h, n, _lib.fptr(samples)),
                   "mf6hip_unet_set_data")
        self.nsamples = n

    def train(self, nsteps, batch, lr, seed=0):
        rng = np.random.default_rng(seed)
        order = np.concatenate([rng.permutation(self.nsamples)
                                for _ in range((nsteps * batch) // self.nsamples + 1)])
        order = np.ascontiguousarray(order[:nsteps * batch], dtype=np.int32)
        loss = np.zeros(nsteps, dtype=np.float32)
        _lib.check(self.L.mf6hip_unet_train(self.h, nsteps, batch, _lib.iptr(order),
                                            float(lr), _lib.fptr(loss)), "mf6hip_unet_train")
        return loss

    def destroy(self):
        if self.h:
            self.L.mf6hip_unet_destroy(self.h)
            self.h = C.c_void_p()


def layer_shapes(nin, ch):
    """(cin, cout) of the 3x3x3 convolutions in parameter order: two per
    level (encoder), then the decoder from the second-coarsest level up."""
    enc = []
    for l, c in enumerate(ch):
        enc += [(nin if l == 0 else ch[l - 1], c), (c, c)]
    dec = [(ch[l + 1] + ch[l], ch[l]) for l in range(len(ch) - 2, -1, -1)]
    return enc + dec


def init_params(nin, ch, rng):
    """He-normal weights, zero biases, in the library's parameter order."""
    out = []
    for a, b in layer_shapes(nin, ch):
        out.append(rng.standard_normal(27 * a * b) * np.sqrt(2.0 / (27 * a)))
        out.append(np.zeros(b))
    out.append(rng.standard_normal(ch[0]) * np.sqrt(1.0 / ch[0]))
    out.append(np.zeros(1))
    return np.concatenate(out).astype(np.float32)


# ---------------------------------------------------------------------------
# systems and training pairs
# ---------------------------------------------------------------------------

@dataclass
class DeviceSystem:
    """The fixed steady system of one case (formulated and row-fixed on the
    device) with a solver handle for solves and single operators."""
    case: object
    amat: np.ndarray
    rhs: np.ndarray
    x0: np.ndarray
    solver: HipLinearSolver
    feat: np.ndarray


def device_system(case, settings: ImsLinearSettings):
    """Formulate the steady system of `case` on the device and return it
    after the sln_ls row fixes (one inner iteration fetches the fixed system)
    with the start heads (CHD heads in place)."""
    import dataclasses
    from .solution import GwfSolution
    c1 = dataclasses.replace(case, ims=dataclasses.replace(case.ims, inner_maximum=1))
    gs = GwfSolution(c1)
    gs.kstp = 1
    gs.prepare_timestep(1.0, True)
    x0 = gs.heads()
    gs.formulate(1)
    gs.linear_solve(1)
    amat, rhs = gs.system()
    gs.destroy()
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    sv = HipLinearSolver().initialize(ia, ja, settings, case.nlay, case.nrow, case.ncol)
    return DeviceSystem(case, amat, rhs, x0, sv, features(case))


def apply_op(sv: HipLinearSolver, op, x):
    y = np.zeros(sv.neq)
    xx = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    _lib.check(sv.L.mf6hip_apply_op(sv.handle, op, _lib.dptr(xx), _lib.dptr(y)),
               "mf6hip_apply_op")
    return y


def training_pairs(sy: DeviceSystem, iterates=(0, 25, 50, 100, 200)):
    """Samples [len(iterates)][4][nodes]: t / s, ln K, ln(K33/K), (e - z) / s
    from FP64 ILU0-PCG iterates x_k (k = iterates; x_0 = the start heads) of
    the device solver."""
    sv = sy.solver
    tight = sv.settings
    xs = sy.x0.copy()
    sv.settings = ImsLinearSettings(**{**tight.__dict__, "dvclose": 1e-11, "rclose": 1e-9,
                                       "iter1": 20000})
    sv.reconfigure()
    sv.solve(1, sy.amat, sy.rhs, xs)
    out = []
    for k in iterates:
        xk = sy.x0.copy()
        if k > 0:
            sv.settings = ImsLinearSettings(**{**tight.__dict__, "dvclose": 0.0,
                                               "rclose": 1e-300, "iter1": int(k)})
            sv.reconfigure()
            sv.solve(1, sy.amat, sy.rhs, xk)
        rk = sy.rhs - apply_op(sv, 0, xk)
        ek = xs - xk
        z = apply_op(sv, 1, rk)
        t = rk - apply_op(sv, 0, z)
        s = np.sqrt(np.mean(t * t))
        if not s > 0:
            continue
        out.append(np.stack([t / s, sy.feat[0], sy.feat[1], (ek - z) / s]).astype(np.float32))
    sv.settings = tight
    sv.reconfigure()
    return out


# ---------------------------------------------------------------------------
# the workload: data, training, evaluation
# ---------------------------------------------------------------------------

def _settings(case, **kw):
    ims = case.ims
    base = dict(dvclose=ims.inner_dvclose, rclose=ims.inner_rclose,
                icnvgopt=ims.icnvgopt, iter1=ims.inner_maximum, ilinmeth=1)
    base.update(kw)
    return ImsLinearSettings(**base)


def make_dataset(seeds, shape=(16, 256, 256), iterates=(0, 25, 50, 100, 200), log=print,
                 per_field=None, pool=None):
    """Training samples from FP64 ILU0-PCG iterates on seeded fields.
    per_field = m takes m of the iterates of each field (a rotation through
    the list by field index), so that many fields fit in memory."""
    out = []
    t0 = time.perf_counter()
    for i, (seed, case) in enumerate(zip(seeds, learned_cases(seeds, shape, pool))):
        sy = device_system(case, _settings(case))
        its = iterates
        if per_field is not None and per_field < len(iterates):
            its = tuple(iterates[(i + q) % len(iterates)] for q in range(per_field))
        out.extend(training_pairs(sy, its))
        sy.solver.destroy()
        if (i + 1) % 8 == 0 or i + 1 == len(seeds):
            log(f"dataset: {i + 1}/{len(seeds)} fields, {len(out)} samples, "
                f"{time.perf_counter() - t0:.0f} s")
    return np.stack(out)


def solve_pair(case, net=None, fcg_m=8, iter1=None, nn_every=1):
    """One steady solve of `case`: FP64 ILU0-PCG (the device fused CG) and
    flexible CG with the network (when given).  Returns a dict of iteration
    counts, device times and head differences."""
    from .solution import GwfSolution
    sy = device_system(case, _settings(case, iter1=iter1 or case.ims.inner_maximum))
    res = {}
    x = sy.x0.copy()
    L = sy.solver.L
    t0 = time.perf_counter()
    ic, it = sy.solver.solve(1, sy.amat, sy.rhs, x)
    res["pcg"] = dict(iterations=it, converged=bool(ic), s=time.perf_counter() - t0)
    xp = x
    sy.solver.destroy()
    sv = HipLinearSolver().initialize(*dis_structure(case.nlay, case.nrow, case.ncol),
                                      _settings(case, preconditioner=1,
                                                iter1=iter1 or case.ims.inner_maximum),
                                      case.nlay, case.nrow, case.ncol)
    periods = [nn_every] if isinstance(nn_every, int) else list(nn_every)
    runs = [("fcg_ilu0", None, 1)]
    if net is not None:
        runs += [("fcg_learned" if k == periods[0] else f"fcg_learned_every{k}", net, k)
                 for k in periods]
    for name, n, k in runs:
        sv.attach_unet(n, sy.feat if n is not None else None, fcg_m, k)
        x = sy.x0.copy()
        t0 = time.perf_counter()
        ic, it = sv.solve(1, sy.amat, sy.rhs, x)
        res[name] = dict(iterations=it, converged=bool(ic), s=time.perf_counter() - t0,
                         learned_iterations=sv.learned_iterations(),
                         max_head_diff_vs_pcg_m=float(np.max(np.abs(x - xp))))
    sv.destroy()
    return res


def run_workload(ntrain=64, ntest=16, shape=(16, 256, 256), epochs=30, batch=4, lr=1e-3,
                 fcg_m=8, scale_up=True, scale_up_iters=20000, nn_every=1, log=print,
                 channels=CHANNELS, per_field=None):
    """Train on `ntrain` seeded fields, evaluate on `ntest` unseen seeds, and
    apply the trained network to the 10 x 1000 x 1000 steady model."""
    t0 = time.perf_counter()
    train_seeds = list(range(1000, 1000 + ntrain))
    test_seeds = list(range(5000, 5000 + ntest))
    pool = field_pool()
    try:
        summary, params = _run_workload(pool, train_seeds, test_seeds, shape, epochs, batch,
                                        lr, fcg_m, scale_up, scale_up_iters, nn_every, log,
                                        channels, per_field, t0)
    finally:
        if pool is not None:
            pool.close()
            pool.join()
    return summary, params


def _run_workload(pool, train_seeds, test_seeds, shape, epochs, batch, lr, fcg_m, scale_up,
                  scale_up_iters, nn_every, log, channels, per_field, t0):
    ntrain, ntest = len(train_seeds), len(test_seeds)
    data = make_dataset(train_seeds, shape, log=log, per_field=per_field, pool=pool)
    net = UNet(*shape, nin=3, channels=channels, seed=0)
    net.set_data(data)
    nsteps = max(1, epochs * data.shape[0] // batch)
    t1 = time.perf_counter()
    losses = []
    chunk = max(1, 64 // batch)
    for s0 in range(0, nsteps, chunk):
        n = min(chunk, nsteps - s0)
        # cosine decay of the learning rate to 5 % over the run
        lr_s = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * s0 / nsteps)))
        losses.extend(net.train(n, batch, lr_s, seed=s0))
        if (s0 // chunk) % 8 == 0:
            log(f"train: step {s0 + n}/{nsteps}, loss {np.mean(losses[-n:]):.4f}, "
                f"{time.perf_counter() - t1:.0f} s")
    train_s = time.perf_counter() - t1
    params = net.get_params()
    evals = []
    for seed, case in zip(test_seeds, learned_cases(test_seeds, shape, pool)):
        r = solve_pair(case, net, fcg_m, nn_every=nn_every)
        evals.append(r)
        log(f"test seed {seed}: PCG {r['pcg']['iterations']} it {r['pcg']['s']:.3f} s, "
            f"FCG+ILU0 {r['fcg_ilu0']['iterations']} it, " + ", ".join(
                f"{k} {v['iterations']} it {v['s']:.3f} s ({v['learned_iterations']} learned)"
                for k, v in r.items() if k.startswith("fcg_learned")))
    net.destroy()
    summary = dict(
        channels=list(channels), levels=len(channels),
        train_fields=ntrain, test_fields=ntest, samples=int(data.shape[0]),
        shape=list(shape), epochs=epochs, batch=batch, train_steps=nsteps, train_s=train_s,
        nn_every=nn_every,
        loss_first=float(np.mean(losses[:10])), loss_last=float(np.mean(losses[-10:])),
        pcg_iterations=float(np.mean([e["pcg"]["iterations"] for e in evals])),
        fcg_ilu0_iterations=float(np.mean([e["fcg_ilu0"]["iterations"] for e in evals])),
        fcg_learned_iterations=float(np.mean([e["fcg_learned"]["iterations"] for e in evals])),
        pcg_s=float(np.mean([e["pcg"]["s"] for e in evals])),
        fcg_ilu0_s=float(np.mean([e["fcg_ilu0"]["s"] for e in evals])),
        fcg_learned_s=float(np.mean([e["fcg_learned"]["s"] for e in evals])),
        max_head_diff_vs_pcg_m=float(max(e["fcg_learned"]["max_head_diff_vs_pcg_m"]
                                         for e in evals)),
        all_converged=all(e["fcg_learned"]["converged"] for e in evals))
    for k in evals[0]:
        if k.startswith("fcg_learned_every"):
            summary[k] = dict(
                iterations=float(np.mean([e[k]["iterations"] for e in evals])),
                s=float(np.mean([e[k]["s"] for e in evals])),
                all_converged=all(e[k]["converged"] for e in evals))
    if scale_up:
        case = configs.steady_10m()
        big = UNet(case.nlay, case.nrow, case.ncol, nin=3, channels=channels, params=params)
        r = solve_pair(case, big, fcg_m, iter1=scale_up_iters, nn_every=nn_every)
        big.destroy()
        summary["scale_up_10m"] = r
        log(f"10M cells: PCG {r['pcg']['iterations']} it {r['pcg']['s']:.2f} s, FCG+net "
            f"{r['fcg_learned']['iterations']} it {r['fcg_learned']['s']:.2f} s")
    summary["total_s"] = time.perf_counter() - t0
    return summary, params
