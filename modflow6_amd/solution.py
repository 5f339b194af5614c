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
iteration.  The NONLINEAR block runs on the device as well: backtracking
(sln_backtracking :2680-2776, ``mf6hip_gwf_backtrack`` +
``mf6hip_gwf_residual_norms``), pseudo-transient continuation for
steady-state Newton models (sln_calc_ptc :2936-2962 / gwf_ptc
gwf.f90:625-679, applied inside ``mf6hip_gwf_linear_solve`` as sln_ls
:2480-2573 does) and SIMPLE / COOLEY / DBD under-relaxation (sln_underrelax
:2989-3114, in ``mf6hip_gwf_outer_update``).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib, binaryfile
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
    nbt: int = 0
    iptc: int = 0


class GwfSolution:
    """One IMS solution holding one DIS GWF model (a GwfCase)."""

    def __init__(self, case, precision=_lib.MF6HIP_FP64, timing=False,
                 inner_fp32_rtol=0.0, inner_fp32_max=0, kernel_variant=0,
                 preconditioner=0, coarse_size=0):
        self.case = c = case
        L = _lib.load()
        self.L = L
        ia, ja = dis_structure(c.nlay, c.nrow, c.ncol)
        self.settings = ImsLinearSettings(
            dvclose=c.ims.inner_dvclose, rclose=c.ims.inner_rclose,
            icnvgopt=c.ims.icnvgopt, iter1=c.ims.inner_maximum,
            ilinmeth=c.ims.ilinmeth, north=c.ims.number_orthogonalizations,
            relax=c.ims.relaxation_factor, precision=precision,
            inner_fp32_rtol=inner_fp32_rtol, inner_fp32_max=inner_fp32_max,
            kernel_variant=kernel_variant, preconditioner=preconditioner,
            coarse_size=coarse_size)
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
        if c.k22 is not None:
            m.k22 = d(c.k22)
        m.istor_coef, m.iconf_ss, m.iorig_ss = (int(c.istor_coef), int(c.iconf_ss),
                                                int(c.iorig_ss))
        if c.thickstrt:
            m.ithickstrt = 1
            m.strt = d(c.strt)
        for nm in ("angle1", "angle2", "angle3"):
            if getattr(c, nm) is not None:
                setattr(m, nm, d(getattr(c, nm)))
        m.icelltype = i(c.icelltype)
        m.ibound = i(np.ones(c.nodes, dtype=np.int32))
        m.inewton = 1 if c.newton else 0
        m.inewtonur = 1 if c.newton_ur else 0
        m.ivarcv = int(c.ivarcv)
        m.idewatcv = int(c.idewatcv)
        m.icellavg = int(c.icellavg)
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
            opt = None
            if p.flowred is not None or p.depth is not None or p.cubic:
                # WEL AUTO_FLOW_REDUCE (the value as read), DRN AUXDEPTHNAME
                # and DEV_CUBIC_SCALING
                opt = _lib.PackageOptionsC()
                if p.flowred is not None:
                    opt.iflowred = 1
                    opt.flowred = float(p.flowred)
                if p.depth is not None:
                    depth = np.ascontiguousarray(p.depth, dtype=np.float64)
                    self._keep.append(depth)
                    opt.idrndepth = 1
                    opt.drndepth = _lib.dptr(depth)
                opt.icubic_scaling = 1 if p.cubic else 0
            _lib.check(L.mf6hip_gwf_set_package_ex(self.h, _lib.PKG[p.kind],
                                                   nodes.size, _lib.iptr(nodes),
                                                   _lib.dptr(v1), _lib.dptr(v2),
                                                   C.byref(opt) if opt is not None else None),
                       f"mf6hip_gwf_set_package_ex({p.kind})")
        if c.hfb is not None:
            self.set_hfb(*c.hfb)
        if c.gnc is not None:
            self.set_gnc(*c.gnc)
        x0 = np.ascontiguousarray(c.strt, dtype=np.float64)
        _lib.check(L.mf6hip_gwf_set_x(self.h, _lib.dptr(x0)), "mf6hip_gwf_set_x")
        if timing:
            _lib.check(L.mf6hip_set_timing(self.h, 1, 16), "mf6hip_set_timing")
        self.kstp = 0
        self.kper = 0
        self.pertim = 0.0
        self.totim = 0.0
        self.iss = 1
        self.delt = 1.0
        self.log: List[List[OuterRecord]] = []
        ims = c.ims
        self.ur = _lib.UnderRelaxC(ims.nonmeth, ims.under_relaxation_gamma,
                                   ims.theta, ims.under_relaxation_kappa,
                                   ims.under_relaxation_momentum, 0.0, 0.0, 0.0)
        self.ptc = _lib.PtcC(0, ims.iallowptc, ims.ptc_option, 0, 0, 0.0,
                             ims.ptc_del0, ims.ptc_exponent, 0.0, 0.0)
        self.res_prev = 0.0
        self.res_new = 0.0

    # gwf_ad + chd_ad
    def prepare_timestep(self, delt, steady):
        self.iss = 1 if steady else 0
        self.delt = float(delt)
        _lib.check(self.L.mf6hip_gwf_advance(self.h, self.iss, self.delt),
                   "mf6hip_gwf_advance")

    def formulate(self, kiter, want_system=False, inewton=1):
        amat = rhs = None
        if want_system:
            amat = np.zeros(self.solver.nja)
            rhs = np.zeros(self.solver.neq)
        _lib.check(self.L.mf6hip_gwf_formulate(self.h, kiter, inewton,
                                               _lib.dptr(amat), _lib.dptr(rhs)),
                   "mf6hip_gwf_formulate")
        return amat, rhs

    def residual_norms(self):
        l2, rv = C.c_double(0.0), C.c_double(0.0)
        _lib.check(self.L.mf6hip_gwf_residual_norms(self.h, C.byref(l2),
                                                    C.byref(rv)),
                   "mf6hip_gwf_residual_norms")
        return l2.value, rv.value

    # sln_backtracking (NumericalSolution.f90:2680-2776)
    def backtracking(self, kiter):
        ims = self.case.ims
        self.formulate(kiter, inewton=0)
        if kiter == 1:
            self.res_prev = self.residual_norms()[0]
        else:
            self.res_new = self.residual_norms()[0]
        nbt = 0
        if kiter > 1:
            if self.res_new > self.res_prev * ims.backtracking_tolerance:
                for nb in range(1, ims.backtracking_number + 1):
                    flag = C.c_int(0)
                    _lib.check(self.L.mf6hip_gwf_backtrack(
                        self.h, ims.backtracking_reduction_factor,
                        ims.outer_dvclose, C.byref(flag)), "mf6hip_gwf_backtrack")
                    if flag.value == 0:
                        break
                    nbt = nb
                    self.formulate(kiter, inewton=0)
                    self.res_new = self.residual_norms()[0]
                    if nb == ims.backtracking_number:
                        break
                    if self.res_new < self.res_prev * ims.backtracking_tolerance:
                        break
                    if self.res_new < ims.backtracking_residual_limit:
                        break
            self.res_prev = self.res_new
        return nbt

    # sln_calc_ptc -> gwf_ptc (gwf.f90:625-679)
    def calc_ptc(self):
        iptc, ptcf = 0, 0.0
        if self.iss > 0 and self.case.newton:
            iptc = 1
            ptcf = self.residual_norms()[1]
            if ptcf == 0.0:
                ptcf = 1.0 / (self.delt * 10.0)
        return iptc, ptcf

    def linear_solve(self, kiter, record=None, iptc=0, ptcf=0.0):
        icnvg, inner, lrch = C.c_int(0), C.c_int(0), C.c_int(0)
        hncg = C.c_double(0.0)
        out = record.to_c() if record is not None else None
        ptc = None
        if iptc:
            self.ptc.iptc = iptc
            self.ptc.ptcf = ptcf
            self.ptc.kstp = self.kstp
            self.ptc.kper = self.kper
            ptc = C.byref(self.ptc)
        _lib.check(self.L.mf6hip_gwf_linear_solve(
            self.h, self.kstp, kiter, ptc, C.byref(icnvg), C.byref(inner),
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
        ims = self.case.ims
        nbt = 0
        if ims.backtracking_number > 0:
            nbt = self.backtracking(kiter)
        self.formulate(kiter)
        iptc, ptcf = self.calc_ptc()
        icnvg, inner, hncg, lrch = self.linear_solve(kiter, record, iptc, ptcf)
        solve_ms = self.last_timing()[0]
        dvclose = ims.outer_dvclose
        conv = abs(hncg) <= dvclose
        rec = OuterRecord(kiter, inner, icnvg, hncg, lrch, conv,
                          solve_ms=solve_ms, nbt=nbt,
                          iptc=self.ptc.iptc if iptc else 0)
        if not conv:
            inur, dxo, h2, l2 = C.c_int(0), C.c_double(0), C.c_double(0), C.c_int(0)
            self.ur.bigch = hncg
            _lib.check(self.L.mf6hip_gwf_outer_update(
                self.h, kiter, C.byref(self.ur), C.byref(inur), C.byref(dxo),
                C.byref(h2), C.byref(l2)), "mf6hip_gwf_outer_update")
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
        self.pertim += delt
        self.totim += delt
        return self.sln_ca()

    def run(self, max_steps=None, head_file=None, budget_file=None, **budget_options):
        """Every stress period and time step of the case (equal steps).
        head_file / budget_file (binary files open for writing, or None)
        receive the heads and the cell-by-cell flows after every step, as
        OC's SAVE HEAD ALL / SAVE BUDGET ALL with SAVE_FLOWS on every
        package write them; budget_options go to budget_records."""
        nsteps = 0
        self.kper = 0
        self.totim = 0.0
        for perlen, nstp, steady in self.case.periods:
            self.kper += 1
            self.kstp = 0
            self.pertim = 0.0
            for istp in range(nstp):
                self.run_timestep(perlen / nstp, steady)
                if istp == nstp - 1:
                    self.pertim = float(perlen)   # end of period (tdis.f90:196-199)
                if budget_file is not None:
                    budget_file.write(self.budget_records(**budget_options))
                if head_file is not None:
                    head_file.write(self.head_records())
                nsteps += 1
                if max_steps is not None and nsteps >= max_steps:
                    return self.heads()
        return self.heads()

    BUDGET_TERMS = ("STO-SS", "STO-SY", "WEL", "DRN", "RCH", "CHD")

    def budget(self, want_flowja=True):
        """gwf_cq + gwf_bd of the current heads on the device: flowja (CSR
        order, positive into the cell, diagonal = flow residual), the budget
        terms {name: (rate in, rate out)} and the percent discrepancy."""
        c = self.case
        nja = c.nodes + 2 * (c.nlay * (c.nrow * (c.ncol - 1) + (c.nrow - 1) * c.ncol)
                             + (c.nlay - 1) * c.nrow * c.ncol)
        fj = np.zeros(nja) if want_flowja else None
        terms = np.zeros(12)
        pct = C.c_double(0.0)
        _lib.check(self.L.mf6hip_gwf_budget(self.h, _lib.dptr(fj), _lib.dptr(terms),
                                            C.byref(pct)), "mf6hip_gwf_budget")
        return dict(flowja=fj, pct_discrepancy=pct.value,
                    terms={k: (terms[2 * t], terms[2 * t + 1])
                           for t, k in enumerate(self.BUDGET_TERMS)})

    def flow_term(self, name):
        """One flow term of the last budget(): STO-SS / STO-SY rates per cell
        (natural order) or a package's simvals in bound order."""
        n = C.c_int(0)
        t = _lib.TERM[name]
        _lib.check(self.L.mf6hip_gwf_flow_term(self.h, t, None, C.byref(n)),
                   "mf6hip_gwf_flow_term")
        out = np.zeros(n.value)
        _lib.check(self.L.mf6hip_gwf_flow_term(self.h, t, _lib.dptr(out), C.byref(n)),
                   "mf6hip_gwf_flow_term")
        return out

    def head_records(self):
        """Head file records (ulasav per layer) of the current heads."""
        c = self.case
        return binaryfile.head_records(self.heads(), c.nlay, c.nrow, c.ncol, self.kstp,
                                       self.kper, self.pertim, self.totim)

    def budget_records(self, model_name="GWF", package_names=None, rch_readasarrays=False):
        """Budget file records (gwf_ot_flow order) of a budget of the current
        heads; package_names maps a package kind to its name (default
        "<KIND>-1", the reference's unnamed-package name)."""
        c = self.case
        b = self.budget(want_flowja=True)
        strgss = strgsy = None
        if c.has_sto:
            strgss = self.flow_term("STO-SS")
            if c.iconvert is not None and np.any(np.asarray(c.iconvert) != 0):
                strgsy = self.flow_term("STO-SY")
        names = package_names or {}
        pk = []
        for p in c.packages:
            kind = "RCHA" if (p.kind == "RCH" and rch_readasarrays) else p.kind
            pk.append((kind, names.get(p.kind, f"{kind}-1"), np.asarray(p.nodes),
                       self.flow_term(p.kind)))
        return binaryfile.budget_records(self.kstp, self.kper, self.delt, self.pertim,
                                         self.totim, c.nlay, c.nrow, c.ncol, b["flowja"],
                                         strgss, strgsy, pk, model_name)

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

    def set_gnc(self, noden, nodem, nodesj, alphasj, implicit=False):
        """The model's GNC package (explicit form), 0-based cells and
        contributing cells (-1: none), (nexg, numjs) weights; replaces the
        last set (mf6hip_gwf_set_gnc)."""
        n = np.ascontiguousarray(np.asarray(noden) + 1, dtype=np.int32)
        m = np.ascontiguousarray(np.asarray(nodem) + 1, dtype=np.int32)
        if n.size == 0:
            _lib.check(self.L.mf6hip_gwf_set_gnc(self.h, 0, 0, None, None, None, None,
                                                 1 if implicit else 0),
                       "mf6hip_gwf_set_gnc")
            return
        js = np.asarray(nodesj).reshape(n.size, -1)
        numjs = js.shape[1]
        j1 = np.ascontiguousarray(js + 1, dtype=np.int32)
        al = np.ascontiguousarray(np.asarray(alphasj, dtype=np.float64).reshape(js.shape))
        _lib.check(self.L.mf6hip_gwf_set_gnc(self.h, n.size, numjs, _lib.iptr(n), _lib.iptr(m),
                                             _lib.iptr(j1), _lib.dptr(al),
                                             1 if implicit else 0),
                   "mf6hip_gwf_set_gnc")

    def set_hfb(self, noden, nodem, hydchr):
        """A new set of horizontal flow barriers (0-based node numbers of
        horizontally adjacent cells, hydraulic characteristics): hfb_rp's
        condsat_reset + condsat_modify (gwf-hfb.f90:149-201, :770-832)."""
        n = np.ascontiguousarray(np.asarray(noden) + 1, dtype=np.int32)
        m = np.ascontiguousarray(np.asarray(nodem) + 1, dtype=np.int32)
        hy = np.ascontiguousarray(hydchr, dtype=np.float64)
        _lib.check(self.L.mf6hip_gwf_set_hfb(self.h, n.size, _lib.iptr(n), _lib.iptr(m),
                                             _lib.dptr(hy)), "mf6hip_gwf_set_hfb")

    def destroy(self):
        self.solver.destroy()


def nonlinear_settings(case):
    """mf6hip_nonlinear of a case's IMS NONLINEAR block."""
    ims = case.ims
    return _lib.NonlinearC(ims.outer_dvclose, ims.outer_maximum, ims.nonmeth,
                           ims.under_relaxation_gamma, ims.under_relaxation_theta,
                           ims.under_relaxation_kappa, ims.under_relaxation_momentum,
                           ims.backtracking_number, ims.backtracking_tolerance,
                           ims.backtracking_reduction_factor,
                           ims.backtracking_residual_limit, ims.iallowptc,
                           ims.ptc_option, ims.ptc_exponent, ims.ptc_del0)


class XmiDriver:
    """A coupler's loop over the XMI-style entry points (srcbmi/mf6xmi.F90
    :220-340): per time step mf6hip_xmi_prepare_solve, mf6hip_xmi_solve
    until it reports convergence or OUTER_MAXIMUM, mf6hip_xmi_finalize_solve.
    The outer iterations run in the library (C++ over the device entry
    points); only scalars reach Python."""

    def __init__(self, gs: GwfSolution):
        self.gs = gs
        self.L = gs.L
        self.nl = nonlinear_settings(gs.case)
        _lib.check(self.L.mf6hip_xmi_set_nonlinear(gs.h, C.byref(self.nl)),
                   "mf6hip_xmi_set_nonlinear")
        self.log: List[List[OuterRecord]] = []

    def run_timestep(self, kper, kstp, delt, steady):
        L, h = self.L, self.gs.h
        _lib.check(L.mf6hip_xmi_prepare_solve(h, kper, kstp, 1 if steady else 0, float(delt)),
                   "mf6hip_xmi_prepare_solve")
        recs = []
        for _ in range(self.nl.outer_maximum):
            hc = C.c_int(0)
            _lib.check(L.mf6hip_xmi_solve(h, C.byref(hc)), "mf6hip_xmi_solve")
            r = _lib.XmiIterationC()
            _lib.check(L.mf6hip_xmi_last_iteration(h, C.byref(r)), "mf6hip_xmi_last_iteration")
            recs.append(OuterRecord(r.kiter, r.inner, r.icnvg_inner, r.hncg, r.lrch,
                                    bool(r.converged), nur=bool(r.nur), nbt=r.nbt,
                                    iptc=r.iptc))
            if hc.value:
                break
        conv, nout = C.c_int(0), C.c_int(0)
        _lib.check(L.mf6hip_xmi_finalize_solve(h, C.byref(conv), C.byref(nout)),
                   "mf6hip_xmi_finalize_solve")
        assert nout.value == len(recs)
        self.log.append(recs)
        return recs

    def run(self):
        """Every stress period and time step of the case (equal steps)."""
        for kper, (perlen, nstp, steady) in enumerate(self.gs.case.periods, start=1):
            for kstp in range(1, nstp + 1):
                self.run_timestep(kper, kstp, perlen / nstp, steady)
        return self.gs.heads()
