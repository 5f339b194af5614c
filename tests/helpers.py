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


def dis_matrix(shape, seed):
    """A symmetric DIS-structured matrix with random off-diagonals (some
    zero, as CHD rows and inactive couplings leave them), in the reference's
    CSR layout."""
    nlay, nrow, ncol = shape
    ia, ja = dis_structure(nlay, nrow, ncol)
    ia0 = ia.astype(np.int64) - 1
    ja0 = ja.astype(np.int64) - 1
    n = nlay * nrow * ncol
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), np.diff(ia0))
    a = np.zeros(ia0[-1])
    upper = ja0 > rows
    a[upper] = rng.uniform(0.1, 3.0, upper.sum()) * (rng.random(upper.sum()) > 0.05)
    # mirror the upper triangle (DIS CSR rows are sorted: diagonal first,
    # then ascending columns)
    key_u = rows[upper] * n + ja0[upper]
    lower = (ja0 < rows)
    key_l = ja0[lower] * n + rows[lower]
    order = np.argsort(key_u)
    a[np.flatnonzero(lower)] = a[np.flatnonzero(upper)[order[np.searchsorted(key_u[order], key_l)]]]
    diag = ia0[:-1]
    a[diag] = -(np.bincount(rows, weights=np.abs(a), minlength=n) + rng.uniform(0.0, 1.0, n))
    return ia, ja, a


def f32_amux(ia, ja, a, x):
    """y = A x in FP32, each row summed in CSR order from 0 (amux order)."""
    ia0 = ia.astype(np.int64) - 1
    ja0 = ja.astype(np.int64) - 1
    n = ia0.size - 1
    a32 = a.astype(np.float32)
    x32 = x.astype(np.float32)
    y = np.zeros(n, dtype=np.float32)
    cnt = np.diff(ia0)
    for s in range(int(cnt.max())):
        m = cnt > s
        e = ia0[:-1][m] + s
        y[m] = (y[m] + a32[e] * x32[ja0[e]]).astype(np.float32)
    return y.astype(np.float64)


def upper_planes(shape, ia, ja, a):
    """The right, front and down couplings of every cell (natural order,
    shape (nlay, nrow, ncol); zero where the neighbour is off the grid) of a
    symmetric DIS matrix in the reference's CSR layout (1-based)."""
    nlay, nrow, ncol = shape
    n = nlay * nrow * ncol
    ia0 = ia.astype(np.int64) - 1
    ja0 = ja.astype(np.int64) - 1
    rows = np.repeat(np.arange(n), np.diff(ia0))
    out = []
    for off in (1, ncol, nrow * ncol):
        m = ja0 == rows + off
        if off == 1:
            m &= (rows % ncol) != ncol - 1
        elif off == ncol:
            m &= ((rows // ncol) % nrow) != nrow - 1
        p = np.zeros(n)
        p[rows[m]] = a[m]
        out.append(p.reshape(shape))
    return out


def rb_ilu_apply(shape, ia, ja, a, piv, r, dtype=np.float64, half_scale_exp=None):
    """z = M^-1 r for the column red-black ILU0 factor of a symmetric DIS
    matrix, colour sweeps in the order of the fused CG's kernels (red
    forward, black forward and backward, red backward), every operation
    rounded to `dtype` (numpy: no fused multiply-add).  piv: the inverse
    pivots (natural order).  half_scale_exp = k: the couplings are stored as
    FP16 values of A 2^-k, the pivots as FP32 values of piv 2^k, and each
    backward sweep starts from its forward value times 2^-k.  Natural-order
    vectors in and out (float64)."""
    nlay, nrow, ncol = shape
    dt = np.dtype(dtype)
    right, front, down = upper_planes(shape, ia, ja, a)
    if half_scale_exp is None:
        A4, A5, A6 = (x.astype(dt) for x in (right, front, down))
        pv = piv.reshape(shape).astype(dt)
        sc = dt.type(1)
    else:
        s = np.ldexp(1.0, -half_scale_exp)
        A4, A5, A6 = ((x * s).astype(np.float16).astype(dt) for x in (right, front, down))
        pv = (piv.reshape(shape) * np.ldexp(1.0, half_scale_exp)).astype(np.float32).astype(dt)
        sc = dt.type(np.ldexp(1.0, -half_scale_exp))
    rr = r.reshape(shape).astype(dt)
    i = np.arange(nrow)[:, None]
    j = np.arange(ncol)[None, :]
    red = ((i + j) & 1) == 0
    hb, hf = i > 0, i < nrow - 1
    hl, hr = j > 0, j < ncol - 1

    def nb(x, di, dj):
        """value of the neighbour (i + di, j + dj) (zero off the grid)"""
        y = np.zeros_like(x)
        ys = slice(max(0, -di), nrow - max(0, di)), slice(max(0, -dj), ncol - max(0, dj))
        xs = slice(max(0, di), nrow + min(0, di)), slice(max(0, dj), ncol + min(0, dj))
        y[ys] = x[xs]
        return y

    zf = np.zeros(shape, dt)
    zup = np.zeros((nrow, ncol), dt)
    lup = np.zeros((nrow, ncol), dt)
    for k in range(nlay):                         # red forward
        tv = rr[k]
        if k > 0:
            tv = tv - lup * zup
        zf[k] = np.where(red, tv, 0)
        zup = tv
        lup = A6[k] * pv[k]
    z = np.zeros(shape, dt)
    for k in range(nlay):                         # black forward
        tb = (nb(A5[k], -1, 0) * nb(pv[k], -1, 0)) * nb(zf[k], -1, 0)
        tl = (nb(A4[k], 0, -1) * nb(pv[k], 0, -1)) * nb(zf[k], 0, -1)
        tr = (A4[k] * nb(pv[k], 0, 1)) * nb(zf[k], 0, 1)
        tf = (A5[k] * nb(pv[k], 1, 0)) * nb(zf[k], 1, 0)
        tv = rr[k]
        tv = np.where(hb, tv - tb, tv)
        tv = np.where(hl, tv - tl, tv)
        tv = np.where(hr, tv - tr, tv)
        tv = np.where(hf, tv - tf, tv)
        if k > 0:
            tv = tv - (A6[k - 1] * pv[k - 1]) * z[k - 1]
        z[k] = np.where(red, 0, tv)
    zdn = np.zeros((nrow, ncol), dt)
    for k in range(nlay - 1, -1, -1):             # black backward
        tv = z[k] * sc
        if k < nlay - 1:
            tv = tv - A6[k] * zdn
        tv = tv * pv[k]
        z[k] = np.where(red, 0, tv)
        zdn = tv
    zdn = np.zeros((nrow, ncol), dt)
    for k in range(nlay - 1, -1, -1):             # red backward
        tv = zf[k] * sc
        if k < nlay - 1:
            tv = tv - A6[k] * zdn
        tv = np.where(hb, tv - nb(A5[k], -1, 0) * nb(z[k], -1, 0), tv)
        tv = np.where(hl, tv - nb(A4[k], 0, -1) * nb(z[k], 0, -1), tv)
        tv = np.where(hr, tv - A4[k] * nb(z[k], 0, 1), tv)
        tv = np.where(hf, tv - A5[k] * nb(z[k], 1, 0), tv)
        tv = tv * pv[k]
        z[k] = np.where(red, tv, z[k])
        zdn = tv
    return z.reshape(-1).astype(np.float64)


def half_scale_exp(shape, ia, ja, a):
    """The power of two the FP16 planes are scaled by: the smallest k with
    max |off-diagonal| 2^-k <= 2^14."""
    amax = max(float(np.max(np.abs(p))) for p in upper_planes(shape, ia, ja, a))
    return int(np.ceil(np.log2(amax / 16384.0))) if amax > 0 else 0


def oracle_rb_factor(shape, ia, ja, a, relax=0.0):
    """pcilu0 (the oracle's) of the column red-black permuted matrix of a
    DIS matrix given in the reference's 1-based CSR: returns (perm, ims
    holding the factor, inverse pivots in natural order, ILU stop row)."""
    perm = O.column_rb_permutation(*shape)
    pia, pja, pa, _ = O.permute_csr(ia.astype(np.int64) - 1, ja.astype(np.int64) - 1,
                                    np.ascontiguousarray(a), perm)
    s = O.ImsSettings()
    s.relax = relax
    ims = O.ImsLinear(pia, pja, s)
    stop = O.lib().orc_pcu(ims.nja, ims.neq, s.ipc, relax, pa, ims.ia, ims.ja, ims.apc,
                           ims.iapc, ims.japc, ims.iwork, ims.work)
    return perm, ims, ims.apc[:ims.neq][perm].copy(), stop


def oracle_rb_ilu0a(perm, ims, r):
    """The oracle's ilu0a of that factor, natural order in and out."""
    pr = np.empty_like(r)
    pr[perm] = r
    pz = np.zeros_like(r)
    O.lib().orc_ilu0a(ims.nja, ims.neq, ims.apc, ims.iapc, ims.japc, pr, pz)
    return pz[perm]


class HipBuffers:
    """Device buffers for tests that hand device pointers to the C-ABI,
    through the HIP runtime the library itself is bound to: after the
    library is loaded, the process's libamdhip64.so.7 (loading the runtime
    by another path would put a second runtime in the process)."""

    def __init__(self):
        _lib.load()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0     # host to device
        self.bufs.append(p)
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []
