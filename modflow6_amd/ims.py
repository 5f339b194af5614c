"""IMS linear solver behind the reference's LinearSolverBaseType interface.

``HipLinearSolver`` mirrors ``LinearSolverBaseType``
(src/Solution/LinearSolverBase.f90:17-61): ``initialize`` (create the device
solver for the CSR structure, :32), ``solve`` (one linear solve, :44, with
``iteration_number`` / ``is_converged`` set as the reference's plugin does),
``print_summary`` and ``destroy``.  It is what a Fortran ``HipLinearSolverType``
binds through ISO_C_BINDING (modflow6_amd/fortran/mf6hip_module.f90).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class ImsLinearSettings:
    """ImsLinearSettingsType (ImsLinearSettings.f90:13-32) plus device
    opt-ins.  Defaults are the reference's init() defaults except where the
    IMS file normally sets a value."""
    dvclose: float = 1e-6
    rclose: float = 1e-6
    icnvgopt: int = 0
    iter1: int = 5000
    ilinmeth: int = 1            # CG_METHOD = 1, BCGS_METHOD = 2
    iscl: int = 0
    iord: int = 0
    north: int = 0
    relax: float = 0.0
    level: int = 0
    droptol: float = 0.0
    precision: int = _lib.MF6HIP_FP64
    ordering: int = _lib.MF6HIP_ORDER_RB
    inner_fp32_max: int = 0
    inner_fp32_rtol: float = 0.0
    kernel_variant: int = 0      # 0 fused CG (symmetric systems), 1 unfused
    preconditioner: int = 0      # 0 ILU0/MILU0, 1 ILU0 + learned correction (FCG),
                                 # 2 two-level (ILU0 + aggregation coarse correction),
                                 # 3 two-level, deflated form
    coarse_size: int = 0         # preconditioner 2, 3: most coarse unknowns
                                 # (0: sqrt(2 N) within [16, 4096])

    def to_c(self, nlay, nrow, ncol):
        return _lib.ImsSettingsC(
            self.dvclose, self.rclose, self.icnvgopt, self.iter1,
            self.ilinmeth, self.iscl, self.iord, self.north, self.relax,
            self.level, self.droptol, self.precision, self.ordering, nlay,
            nrow, ncol, self.inner_fp32_max, self.inner_fp32_rtol,
            self.kernel_variant, self.preconditioner, self.coarse_size)


def dis_structure(nlay, nrow, ncol):
    """1-based CSR structure of a full DIS grid, diagonal first then
    ascending columns (Connections.f90:523-603, Sparse.f90:217-239)."""
    nodes = nlay * nrow * ncol
    nrc = nrow * ncol
    n = np.arange(nodes, dtype=np.int64)
    k = n // nrc
    i = (n // ncol) % nrow
    j = n % ncol
    offs = [(-nrc, k > 0), (-ncol, i > 0), (-1, j > 0), (1, j < ncol - 1),
            (ncol, i < nrow - 1), (nrc, k < nlay - 1)]
    cnt = np.ones(nodes, dtype=np.int64)
    for _, m in offs:
        cnt += m
    ia = np.zeros(nodes + 1, dtype=np.int64)
    ia[1:] = np.cumsum(cnt)
    ja = np.empty(int(ia[-1]), dtype=np.int64)
    pos = ia[:-1].copy()
    ja[pos] = n
    pos += 1
    for off, m in offs:
        idx = np.nonzero(m)[0]
        ja[pos[idx]] = n[idx] + off
        pos[idx] += 1
    return (np.ascontiguousarray(ia + 1, dtype=np.int32),
            np.ascontiguousarray(ja + 1, dtype=np.int32))


class ConvergenceRecord:
    """Per-inner-iteration values (ConvergenceSummaryType fields)."""

    def __init__(self, cap):
        self.cap = cap
        self.dvmax = np.zeros(cap)
        self.rmax = np.zeros(cap)
        self.alpha = np.zeros(cap)
        self.omega = np.zeros(cap)
        self.locdv = np.zeros(cap, dtype=np.int32)
        self.locr = np.zeros(cap, dtype=np.int32)
        self.count = 0

    def to_c(self):
        return _lib.CnvgOutC(self.cap, 0, _lib.dptr(self.dvmax),
                             _lib.dptr(self.rmax), _lib.dptr(self.alpha),
                             _lib.dptr(self.omega), _lib.iptr(self.locdv),
                             _lib.iptr(self.locr))


class HipLinearSolver:
    """LinearSolverBaseType implemented on the GPU (libmf6hip)."""

    def __init__(self, name="SLN_1"):
        self.name = name
        self.iteration_number = 0
        self.is_converged = 0
        self.handle = C.c_void_p()
        self.nja = 0
        self.neq = 0

    # LinearSolverBase.f90:32 initialize(matrix, linear_settings, summary)
    def initialize(self, ia, ja, settings: ImsLinearSettings, nlay, nrow,
                   ncol):
        L = _lib.load()
        self.L = L
        self.neq = ia.size - 1
        self.nja = ja.size
        self.settings = settings
        self._structure = (ia, ja, nlay, nrow, ncol)
        s = settings.to_c(nlay, nrow, ncol)
        _lib.check(L.mf6hip_create(self.neq, self.nja, _lib.iptr(ia),
                                   _lib.iptr(ja), C.byref(s),
                                   C.byref(self.handle)), "mf6hip_create")
        return self

    # LinearSolverBase.f90:44 solve(kiter, rhs, x, cnvg_summary)
    def solve(self, kiter, amat, rhs, x, kstp=1, record=None):
        L = _lib.load()
        icnvg = C.c_int(0)
        inner = C.c_int(0)
        out = record.to_c() if record is not None else None
        _lib.check(L.mf6hip_solve(self.handle, kstp, kiter, _lib.dptr(amat),
                                  _lib.dptr(rhs), _lib.dptr(x),
                                  C.byref(icnvg), C.byref(inner),
                                  C.byref(out) if out is not None else None),
                   "mf6hip_solve")
        if record is not None:
            record.count = out.count
        self.iteration_number = inner.value
        self.is_converged = icnvg.value
        return icnvg.value, inner.value

    def reconfigure(self):
        """Recreate the handle with self.settings (same structure)."""
        self.destroy()
        ia, ja, nlay, nrow, ncol = self._structure
        return self.initialize(ia, ja, self.settings, nlay, nrow, ncol)

    def attach_unet(self, net, feat, fcg_m=8, nn_every=1):
        """Attach a trained network (learned.UNet) with its two feature
        planes; used when settings.preconditioner == 1."""
        f = None if feat is None else np.ascontiguousarray(feat, dtype=np.float32)
        _lib.check(self.L.mf6hip_attach_unet(self.handle, net.h if net else None,
                                             _lib.fptr(f), int(fcg_m), int(nn_every)),
                   "mf6hip_attach_unet")

    def learned_iterations(self):
        n = C.c_int(0)
        _lib.check(self.L.mf6hip_last_learned_iterations(self.handle, C.byref(n)),
                   "mf6hip_last_learned_iterations")
        return n.value

    def print_summary(self):
        return (f"{self.name}: HIP IMS linear solver, {self.neq} equations, "
                f"{self.nja} non-zeros")

    def destroy(self):
        if self.handle:
            _lib.load().mf6hip_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
