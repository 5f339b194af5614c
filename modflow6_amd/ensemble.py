"""Ensembles of independent steady GWF realisations solved together on the
device (``mf6hip_batch_*``): every realisation is formulated on the device
(the same formulate as a single model), its fixed system is handed to the
ensemble, and all of them are solved by one mixed-precision PCG whose
kernel launches cover every realisation, replayed as a HIP graph.

The workload: realisations of uneven size (nlay on [2, 8], nrow = ncol on
[64, 512]), each with its own iid lognormal K (ln-mean ln 10, sigma on
[1, 2]), CHD 10 / 0 m on the first / last column, RCH 1e-4 m/d and 0 to 50
wells at -500 m3/d (the small steady model's recipe).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import time

import numpy as np

from . import _lib, configs, fields
from .configs import GwfCase, ImsOptions
from .ims import ImsLinearSettings, dis_structure


def realisation(seed, nlay_range=(2, 8), n_range=(64, 512), sigma_range=(1.0, 2.0),
                nwel_range=(0, 50)):
    rng = np.random.default_rng(seed)
    nlay = int(rng.integers(nlay_range[0], nlay_range[1] + 1))
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    sigma = float(rng.uniform(*sigma_range))
    nwel = int(rng.integers(nwel_range[0], nwel_range[1] + 1))
    nodes = nlay * n * n
    k11 = fields.iid_lognormal(nodes, np.log(10.0), sigma, rng)
    botm = np.stack([np.full((n, n), -10.0 * (k + 1)) for k in range(nlay)])
    pk = [configs._rch_top(n, n, 1e-4), configs._chd_columns(nlay, n, n, 10.0, 0.0)]
    if nwel > 0:
        pk.insert(0, configs._wells(rng, nlay, n, n, nwel, -500.0))
    return GwfCase(
        name=f"realisation_{seed}_{nlay}x{n}x{n}", nlay=nlay, nrow=n, ncol=n,
        delr=np.full(n, 100.0), delc=np.full(n, 100.0), top2d=np.zeros((n, n)),
        botm=botm, k11=k11, k33=0.1 * k11, icelltype=np.zeros(nodes, dtype=np.int32),
        strt=np.zeros(nodes), packages=pk, ims=ImsOptions(inner_maximum=20000),
        periods=[(1.0, 1, True)], meta=dict(seed=seed, sigma=sigma, nwel=nwel))


def fixed_system(case):
    """(amat, rhs, x0) of the first outer iteration after the sln_ls row
    fixes, formulated on the device."""
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
    return amat, rhs, x0


class Ensemble:
    """ctypes handle of one device ensemble."""

    def __init__(self, cases, log=None, keep_systems=False, preconditioner=0,
                 systems=None):
        """preconditioner 0: ILU0; 2: ILU0 plus the aggregation coarse
        correction (default coarse size per realisation).  systems: the
        (amat, rhs, x0) of each realisation when already formulated."""
        self.L = L = _lib.load()
        self.cases = list(cases)
        self.systems = [] if keep_systems else None
        self.h = C.c_void_p()
        _lib.check(L.mf6hip_batch_create(len(self.cases), C.byref(self.h)),
                   "mf6hip_batch_create")
        for c in self.cases:
            s = ImsLinearSettings(dvclose=c.ims.inner_dvclose, rclose=c.ims.inner_rclose,
                                  iter1=c.ims.inner_maximum, ilinmeth=1,
                                  precision=_lib.MF6HIP_MIXED,
                                  preconditioner=preconditioner)
            cs = s.to_c(c.nlay, c.nrow, c.ncol)
            _lib.check(L.mf6hip_batch_add(self.h, C.byref(cs)), "mf6hip_batch_add")
        _lib.check(L.mf6hip_batch_finalize(self.h), "mf6hip_batch_finalize")
        t0 = time.perf_counter()
        for i, c in enumerate(self.cases):
            amat, rhs, x0 = systems[i] if systems is not None else fixed_system(c)
            if self.systems is not None:
                self.systems.append((amat, rhs, x0))
            ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
            _lib.check(L.mf6hip_batch_set_system(self.h, i, _lib.iptr(ia), _lib.iptr(ja),
                                                 amat.size, _lib.dptr(amat), _lib.dptr(rhs),
                                                 _lib.dptr(x0)), "mf6hip_batch_set_system")
            if log and ((i + 1) % 32 == 0 or i + 1 == len(self.cases)):
                log(f"ensemble: {i + 1}/{len(self.cases)} systems staged, "
                    f"{time.perf_counter() - t0:.0f} s")

    def autotune(self):
        buf = C.create_string_buffer(4096)
        _lib.check(self.L.mf6hip_batch_autotune(self.h, buf, 4096), "mf6hip_batch_autotune")
        return buf.value.decode()

    def solve(self):
        n = len(self.cases)
        icnvg = np.zeros(n, dtype=np.int32)
        iters = np.zeros(n, dtype=np.int32)
        ms = C.c_double(0.0)
        _lib.check(self.L.mf6hip_batch_solve(self.h, _lib.iptr(icnvg), _lib.iptr(iters),
                                             C.byref(ms)), "mf6hip_batch_solve")
        return icnvg, iters, ms.value

    def heads(self, i):
        x = np.zeros(self.cases[i].nodes)
        _lib.check(self.L.mf6hip_batch_get_x(self.h, i, _lib.dptr(x)), "mf6hip_batch_get_x")
        return x

    def destroy(self):
        if self.h:
            self.L.mf6hip_batch_destroy(self.h)
            self.h = C.c_void_p()


def coarse_size_for(case):
    """Coarse size of the two-level preconditioner for one realisation: the
    library's default, sqrt(2 N) within [16, 4096]."""
    return 0


def solve_one_by_one(cases, systems, log=None, precision=_lib.MF6HIP_MIXED,
                     preconditioner=0, keep=0):
    """The same realisations solved one at a time by the single-model device
    solver (the fused CG, one handle per realisation, created outside the
    timing; preconditioner 2 with coarse_size_for each realisation).
    Returns (seconds summed over the solves, iterations, heads of the first
    `keep` realisations)."""
    from .ims import HipLinearSolver
    total, its, heads = 0.0, [], []
    for i, (c, (amat, rhs, x0)) in enumerate(zip(cases, systems)):
        s = ImsLinearSettings(dvclose=c.ims.inner_dvclose, rclose=c.ims.inner_rclose,
                              iter1=c.ims.inner_maximum, ilinmeth=1,
                              precision=precision, preconditioner=preconditioner,
                              coarse_size=coarse_size_for(c))
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        sv = HipLinearSolver().initialize(ia, ja, s, c.nlay, c.nrow, c.ncol)
        x = x0.copy()
        sv.solve(1, amat, rhs, x)           # warm: first call builds the factor path
        x = x0.copy()
        t0 = time.perf_counter()
        _, it = sv.solve(1, amat, rhs, x)
        total += time.perf_counter() - t0
        its.append(it)
        if i < keep:
            heads.append(x.copy())
        sv.destroy()
        if log and ((i + 1) % 64 == 0 or i + 1 == len(cases)):
            log(f"one by one: {i + 1}/{len(cases)} solved, {total:.2f} s of solves")
    return total, its, heads
