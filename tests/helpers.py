"""Shared helpers for the parity tests (oracle systems, device handles)."""
from __future__ import annotations

import ctypes as C
import copy

import numpy as np

from modflow6_amd import _lib, configs
from modflow6_amd.ims import HipLinearSolver, ImsLinearSettings, dis_structure
from oracle import oracle as O


def oracle_system(case, kiter=1, steps_before=0):
    """Oracle GwfSimulation advanced to the first outer iteration of a step
    and its system after formulate + sln_ls row fixes.  Returns
    (sim, amat, rhs, x0) with x0 the heads the solve starts from."""
    sim = O.GwfSimulation(case)
    for _ in range(steps_before):
        perlen, nstp, steady = case.periods[0]
        sim.run_timestep(perlen / nstp, steady)
    sim.kstp += 1
    perlen, nstp, steady = case.periods[0]
    sim.advance(perlen / nstp, steady)
    sim.formulate(kiter)
    d = sim.dis
    O.lib().orc_sln_ls_fixes(sim.neq, d.ia, d.ja, sim.ibound, sim.isymmetric,
                             sim.amat, sim.rhs, sim.x, sim.xtemp)
    return sim, sim.amat.copy(), sim.rhs.copy(), sim.x.copy()


def device_solver(case, precision=_lib.MF6HIP_FP64, **kw):
    ims = case.ims
    s = ImsLinearSettings(dvclose=ims.inner_dvclose, rclose=ims.inner_rclose,
                          icnvgopt=ims.icnvgopt, iter1=ims.inner_maximum,
                          ilinmeth=ims.ilinmeth,
                          north=ims.number_orthogonalizations,
                          relax=ims.relaxation_factor, precision=precision,
                          **kw)
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    return HipLinearSolver().initialize(ia, ja, s, case.nlay, case.nrow,
                                        case.ncol)


def apply_op(solver, op, x=None):
    y = np.zeros(solver.neq)
    xx = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    _lib.check(_lib.load().mf6hip_apply_op(solver.handle, op, _lib.dptr(xx),
                                           _lib.dptr(y)), "mf6hip_apply_op")
    return y


def load_factor(solver, amat):
    _lib.check(_lib.load().mf6hip_load_factor(
        solver.handle, _lib.dptr(np.ascontiguousarray(amat))),
        "mf6hip_load_factor")


def oracle_permuted_factor(sim, amat, relax=0.0):
    """pcilu0 of the column red-black permuted matrix: returns (perm,
    permuted CSR (ia, ja, a), apc, iapc, japc)."""
    d = sim.dis
    perm = O.column_rb_permutation(d.nlay, d.nrow, d.ncol)
    pia, pja, pa, _ = O.permute_csr(d.ia.astype(np.int64),
                                    d.ja.astype(np.int64), amat, perm)
    s = copy.copy(sim.settings)
    s.relax = relax
    ims = O.ImsLinear(pia, pja, s)
    stop = O.lib().orc_pcu(ims.nja, ims.neq, s.ipc, relax, pa, ims.ia, ims.ja,
                           ims.apc, ims.iapc, ims.japc, ims.iwork, ims.work)
    return perm, (pia, pja, pa), ims, stop


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64),
                                                 b.view(np.int64))


def ulp_diff(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    return int(np.max(np.abs(a - b))) if a.size else 0
