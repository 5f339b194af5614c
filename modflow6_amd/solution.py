"""NumericalSolution outer loop over the device-resident GWF system.

Mirrors ``NumericalSolutionType`` (src/Solution/NumericalSolution.f90) for one
structured GWF model whose matrix fill and linear solve run on the GPU:

* ``prepare_timestep`` -- gwf_ad (gwf.f90:396-421) + chd_ad: ``mf6hip_gwf_advance``
* ``solve(kiter)``     -- sln_buildsystem (:1941) ``mf6hip_gwf_formulate``,
                          sln_ls (:2404) + sln_get_dxmax (:3122)
                          ``mf6hip_gwf_linear_solve``, convergence test
                          (:1625-1629), sln_calcdx / npf_nur (:1722-1760)
                          ``mf6hip_gwf_outer_update``
* ``sln_ca``           -- outer iterations until |hncg| <= OUTER_DVCLOSE
                          (:1287-1327)

Heads stay in HBM between calls; only scalars cross to the host each outer
iteration.  Under-relaxation (NONLINEAR UNDER_RELAXATION), backtracking and
pseudo-transient continuation are not part of the device path and are
rejected.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib
from .ims import HipLinearSolver, ImsLinearSettings, dis_structure


@dataclass
class OuterRecord:
    kiter: int
    inner: int
    icnvg_inner: int
    hncg: float
    lrch: int
    converged: bool
    nur: bool = False
    solve_ms: float = 0.0


class GwfSolution:
    """One IMS solution holding one DIS GWF model (a GwfCase)."""

    def __init__(self, case, precision=_lib.MF6HIP_FP64, timing=False,
                 inner_fp32_rtol=0.0, inner_fp32_max=0, kernel_variant=0):
        self.case = c = case
        if c.ims.under_relaxation.upper() != "NONE":
            raise _lib.Mf6HipError("UNDER_RELAXATION is not supported on the "
                                   "device path")
        L = _lib.load()
        self.L = L
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        self.settings = ImsLinearSettings(
            dvclose=c.ims.inner_dvclose, rclose=c.ims.inner_rclose,
            icnvgopt=c.ims.icnvgopt, iter1=c.ims.inner_maximum,
            ilinmeth=c.ims.ilinmeth, north=c.ims.number_orthogonalizations,
            relax=c.ims.relaxation_factor, precision=precision,
            inner_fp32_rtol=inner_fp32_rtol, inner_fp32_max=inner_fp32_max,
            kernel_variant=kernel_variant)
        self.solver = HipLinearSolver().initialize(ia, ja, self.settings,
                                                   c.nlay, c.nrow, c.ncol)
        self.h = self.solver.handle
        top, bot = c.cell_top_bot()
        self._keep = []   # arrays referenced by the model struct

        def d(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _lib.dptr(a)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return _lib.iptr(a)

        m = _lib.GwfModelC()
        m.nlay, m.nrow, m.ncol = c.nlay, c.nrow, c.ncol
        m.delr, m.delc = d(c.delr), d(c.delc)
        m.top, m.bot = d(top), d(bot)
        m.k11, m.k33 = d(c.k11), d(c.k33)
        m.icelltype = i(c.icelltype)
        m.ibound = i(np.ones(c.nodes, dtype=np.int32))
        m.inewton = 1 if c.newton else 0
        m.inewtonur = 1 if c.newton_ur else 0
        m.ivarcv = 0
        m.idewatcv = 0
        m.icellavg = 0
        m.has_sto = 1 if c.has_sto else 0
        if c.has_sto:
            m.iconvert, m.ss, m.sy = i(c.iconvert), d(c.ss), d(c.sy)
        _lib.check(L.mf6hip_gwf_create(self.h, C.byref(m)), "mf6hip_gwf_create")
        self._keep = []
        for p in c.packages:
            nodes = np.ascontiguousarray(np.asarray(p.nodes) + 1, dtype=np.int32)
            v1 = np.ascontiguousarray(p.v1, dtype=np.float64)
            v2 = (np.ascontiguousarray(p.v2, dtype=np.float64)
                  if p.v2 is not None else None)
            _lib.check(L.mf6hip_gwf_set_package(self.h, _lib.PKG[p.kind],
                                                nodes.size, _lib.iptr(nodes),
                                                _lib.dptr(v1), _lib.dptr(v2)),
                       f"mf6hip_gwf_set_package({p.kind})")
        x0 = np.ascontiguousarray(c.strt, dtype=np.float64)
        _lib.check(L.mf6hip_gwf_set_x(self.h, _lib.dptr(x0)), "mf6hip_gwf_set_x")
        if timing:
            _lib.check(L.mf6hip_set_timing(self.h, 1, 16), "mf6hip_set_timing")
        self.kstp = 0
        self.log: List[List[OuterRecord]] = []

    # gwf_ad + chd_ad
    def prepare_timestep(self, delt, steady):
        _lib.check(self.L.mf6hip_gwf_advance(self.h, 1 if steady else 0,
                                             float(delt)),
                   "mf6hip_gwf_advance")

    def formulate(self, kiter, want_system=False):
        amat = rhs = None
        if want_system:
            amat = np.zeros(self.solver.nja)
            rhs = np.zeros(self.solver.neq)
        _lib.check(self.L.mf6hip_gwf_formulate(self.h, kiter, _lib.dptr(amat),
                                               _lib.dptr(rhs)),
                   "mf6hip_gwf_formulate")
        return amat, rhs

    def linear_solve(self, kiter, record=None):
        icnvg, inner, lrch = C.c_int(0), C.c_int(0), C.c_int(0)
        hncg = C.c_double(0.0)
        out = record.to_c() if record is not None else None
        _lib.check(self.L.mf6hip_gwf_linear_solve(
            self.h, self.kstp, kiter, C.byref(icnvg), C.byref(inner),
            C.byref(hncg), C.byref(lrch),
            C.byref(out) if out is not None else None),
            "mf6hip_gwf_linear_solve")
        if record is not None:
            record.count = out.count
        return icnvg.value, inner.value, hncg.value, lrch.value

    def last_timing(self):
        ms, spmv, n = C.c_double(0), C.c_double(0), C.c_int(0)
        _lib.check(self.L.mf6hip_last_timing(self.h, C.byref(ms),
                                             C.byref(spmv), C.byref(n)),
                   "mf6hip_last_timing")
        return ms.value, spmv.value, n.value

    # NumericalSolution%solve (:1482-1837)
    def solve(self, kiter, record=None):
        self.formulate(kiter)
        icnvg, inner, hncg, lrch = self.linear_solve(kiter, record)
        solve_ms = self.last_timing()[0]
        dvclose = self.case.ims.outer_dvclose
        conv = abs(hncg) <= dvclose
        rec = OuterRecord(kiter, inner, icnvg, hncg, lrch, conv,
                          solve_ms=solve_ms)
        if not conv:
            inur, dxo, h2, l2 = C.c_int(0), C.c_double(0), C.c_double(0), C.c_int(0)
            _lib.check(self.L.mf6hip_gwf_outer_update(
                self.h, kiter, C.byref(inur), C.byref(dxo), C.byref(h2),
                C.byref(l2)), "mf6hip_gwf_outer_update")
            if inur.value:
                rec.nur = True
                # sln_nur_has_converged (:3205-3218)
                if abs(dxo.value) <= dvclose and abs(hncg) <= dvclose:
                    rec.converged = True
                    rec.hncg, rec.lrch = h2.value, l2.value
        return rec

    def sln_ca(self):
        recs = []
        for kiter in range(1, self.case.ims.outer_maximum + 1):
            rec = self.solve(kiter)
            recs.append(rec)
            if rec.converged:
                break
        self.log.append(recs)
        return recs

    def run_timestep(self, delt, steady):
        self.kstp += 1
        self.prepare_timestep(delt, steady)
        return self.sln_ca()

    def run(self, max_steps=None):
        nsteps = 0
        for perlen, nstp, steady in self.case.periods:
            self.kstp = 0
            for _ in range(nstp):
                self.run_timestep(perlen / nstp, steady)
                nsteps += 1
                if max_steps is not None and nsteps >= max_steps:
                    return self.heads()
        return self.heads()

    def heads(self):
        x = np.zeros(self.case.nodes)
        _lib.check(self.L.mf6hip_gwf_get_x(self.h, _lib.dptr(x)),
                   "mf6hip_gwf_get_x")
        return x

    def system(self):
        """(amat, rhs) of the last linear solve after the sln_ls row fixes,
        reference CSR layout."""
        amat = np.zeros(self.solver.nja)
        rhs = np.zeros(self.solver.neq)
        _lib.check(self.L.mf6hip_get_system(self.h, _lib.dptr(amat),
                                            _lib.dptr(rhs)),
                   "mf6hip_get_system")
        return amat, rhs

    def destroy(self):
        self.solver.destroy()
