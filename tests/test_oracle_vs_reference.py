"""Pins the C oracle (oracle/mf6_oracle.c) to the reference's own Fortran.

oracle/_ref/libmf6ref.so is the reference's IMS core and GWF utility
modules (ImsLinearBase.f90, GwfConductanceUtils.f90, SmoothingFunctions.f90,
GwfStorageUtils.f90, sparsekit amux, blas ddot / dnrm2, ...) compiled from
/root/reference by oracle/Makefile; every comparison here is bit for bit.
Skipped when that library has not been built (it needs the reference
sources)."""
import ctypes as C

import numpy as np
import pytest

from modflow6_amd import configs
from oracle import oracle as O
from tests.helpers import oracle_system

R = O.ref_lib()
L = O.lib()
pytestmark = pytest.mark.skipif(R is None, reason="oracle/_ref not built")


def eqbits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.int64),
                          np.asarray(b, dtype=np.float64).view(np.int64))


def test_blas_sparsekit():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 1000, 100003):
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
        y = rng.standard_normal(n)
        assert eqbits(L.orc_ddot(n, x, y), R.ref_ddot(n, x, y))
        assert eqbits(L.orc_dnrm2(n, x), R.ref_dnrm2(n, x))
    case = configs.steady_small(nlay=2, nrow=12, ncol=9, nwel=4)
    sim, amat, rhs, x0 = oracle_system(case)
    d = sim.dis
    x = rng.standard_normal(sim.neq)
    y1, y2 = np.zeros(sim.neq), np.zeros(sim.neq)
    L.orc_amux(sim.neq, x, y1, amat, d.ja, d.ia)
    R.ref_amux(sim.neq, x, y2, amat, (d.ja + 1).astype(np.int32),
               (d.ia + 1).astype(np.int32))
    assert eqbits(y1, y2)
    r1, r2 = np.zeros(sim.neq), np.zeros(sim.neq)
    L.orc_residual(sim.neq, x, rhs, r1, amat, d.ia, d.ja)
    R.ref_residual(sim.neq, d.nja, x, rhs, r2, amat, (d.ia + 1).astype(np.int32),
                   (d.ja + 1).astype(np.int32))
    assert eqbits(r1, r2)


def test_scalar_helpers():
    rng = np.random.default_rng(1)
    for _ in range(4000):
        top = rng.uniform(-5, 50)
        bot = top - rng.choice([0.0, rng.uniform(0.1, 20)])
        x = rng.uniform(bot - 3, top + 3) if rng.random() < 0.8 else \
            rng.choice([bot, top, bot + 1e-7 * (top - bot), top - 1e-7 * (top - bot)])
        for eps in (0.0, 1e-6, 1e-3):
            assert eqbits(L.orc_sqsat(top, bot, x, eps), R.ref_sqsat(top, bot, x, eps))
            if top > bot:
                assert eqbits(L.orc_sqsat_deriv(top, bot, x, eps),
                              R.ref_sqsat_deriv(top, bot, x, eps))
        # sQSaturation / sQSaturationDerivative (cubic smoothing of WEL
        # AUTO_FLOW_REDUCE and DRN depth scaling), both coefficient pairs
        if top > bot:
            for c1, c2 in ((-2.0, 3.0), (-1.0, 2.0)):
                assert eqbits(L.orc_sqsat_cubic(top, bot, x, c1, c2),
                              R.ref_sqsat_cubic(top, bot, x, c1, c2))
                assert eqbits(L.orc_sqsat_cubic_deriv(top, bot, x, c1, c2),
                              R.ref_sqsat_cubic_deriv(top, bot, x, c1, c2))
    for _ in range(3000):
        k1, k2 = rng.lognormal(0, 2, 2) * rng.choice([1, 1, 0], 2)
        t1, t2 = rng.uniform(0, 20, 2)
        cl1, cl2, w = rng.uniform(1, 100, 3)
        for iavg in range(4):
            assert eqbits(L.orc_condmean(k1, k2, t1, t2, cl1, cl2, w, iavg),
                          R.ref_condmean(k1, k2, t1, t2, cl1, cl2, w, iavg))
        ib = rng.integers(0, 2, 2)
        ic = rng.integers(0, 2, 2)
        ups = int(rng.integers(0, 2))
        hn, hm = rng.uniform(-10, 30, 2)
        sn, sm = rng.uniform(0, 1, 2)
        top_n, top_m = rng.uniform(10, 30, 2)
        bot_n, bot_m = top_n - rng.uniform(1, 10), top_m - rng.uniform(1, 10)
        iavg = int(rng.integers(0, 4))
        args = (int(ib[0]), int(ib[1]), int(ic[0]), int(ic[1]), ups, 1, iavg,
                float(rng.lognormal()), hn, hm, sn, sm, k1 + 0.1, k2 + 0.1, top_n,
                top_m, bot_n, bot_m, cl1, cl2, w)
        assert eqbits(L.orc_hcond(*args), R.ref_hcond(*args))
        ivar, idw = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        vargs = (int(ib[0]), int(ib[1]), int(ic[0]), int(ic[1]), 0, ivar, idw,
                 float(rng.lognormal()), hn, hm, k1 + 0.1, k2 + 0.1, sn, sm, top_n,
                 top_m, bot_n, bot_m, w)
        assert eqbits(L.orc_vcond(*vargs), R.ref_vcond(*vargs))
    for _ in range(2000):
        top = rng.uniform(0, 50)
        bot = top - rng.uniform(0.5, 20)
        area, ss, sy = rng.uniform(1, 1e4), rng.lognormal(-11, 1), rng.uniform(0.01, 0.3)
        for ic in (0, 1):
            assert eqbits(L.orc_sscapacity(ic, top, bot, area, ss),
                          R.ref_sscapacity(ic, top, bot, area, ss))
        assert eqbits(L.orc_sycapacity(area, sy), R.ref_sycapacity(area, sy))
        rho1, rho1o, rho2, rho2o = rng.lognormal(0, 1, 4)
        snn, sno = rng.choice([0.0, 1.0, rng.uniform()], 2)
        hn, ho = rng.uniform(bot - 2, top + 2, 2)
        for iconv in (0, 1):
            for iorig in (0, 1):
                for iconf in (0, 1):
                    a1, b1, a2, b2 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
                    L.orc_ssterms(iconv, iorig, iconf, top, bot, rho1, rho1o, snn, sno,
                                  hn, ho, C.byref(a1), C.byref(b1))
                    R.ref_ssterms(iconv, iorig, iconf, top, bot, rho1, rho1o, snn, sno,
                                  hn, ho, C.byref(a2), C.byref(b2))
                    assert eqbits(a1.value, a2.value) and eqbits(b1.value, b2.value)
        a1, b1, a2, b2 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        L.orc_syterms(top, bot, rho2, rho2o, snn, sno, C.byref(a1), C.byref(b1))
        R.ref_syterms(top, bot, rho2, rho2o, snn, sno, C.byref(a2), C.byref(b2))
        assert eqbits(a1.value, a2.value) and eqbits(b1.value, b2.value)


def test_convergence_helpers():
    rng = np.random.default_rng(2)
    for _ in range(3000):
        a = rng.standard_normal()
        b = a * (1 + rng.choice([0, 1e-15, 1e-14, 1e-13, 1e-10]) * rng.choice([-1, 1]))
        assert L.orc_is_close(a, b) == R.ref_is_close(a, b)
        args = [float(v) for v in (rng.choice([1e-7, 1e-5, 1e-3]) * rng.standard_normal(),
                                   rng.choice([1e-7, 1e-5, 1e-3]) * rng.standard_normal(),
                                   rng.uniform(0, 1), rng.uniform(0, 0.2), 1e-6, 1e-4)]
        for opt in range(5):
            for it in (1, 2):
                i1, i2 = C.c_int(0), C.c_int(0)
                L.orc_testcnvg(opt, C.byref(i1), it, *args)
                R.ref_testcnvg(opt, C.byref(i2), it, *args)
                assert i1.value == i2.value
    for opt in range(5):
        for kstp in (1, 2):
            assert eqbits(L.orc_epfact(opt, kstp), R.ref_epfact(opt, kstp))


def _ref_factor(sim, amat, ipc, relax):
    d = sim.dis
    ia1 = (d.ia + 1).astype(np.int32)
    ja1 = (d.ja + 1).astype(np.int32)
    iapc = np.zeros(sim.neq + 1, np.int32)
    japc = np.zeros(d.nja, np.int32)
    apc = np.zeros(d.nja)
    R.ref_pccrs(sim.neq, d.nja, ia1, ja1, iapc, japc)
    R.ref_pcu(sim.neq, d.nja, ipc, relax, amat, ia1, ja1, apc, iapc, japc)
    return apc, iapc, japc


@pytest.mark.parametrize("relax", [0.0, 0.97])
@pytest.mark.parametrize("which", ["confined", "newton"])
def test_ilu0_factor_apply(relax, which):
    if which == "confined":
        case = configs.steady_small(nlay=3, nrow=20, ncol=16, nwel=6)
    else:
        case = configs.transient_newton(seed=4, nlay=3, nrow=14, ncol=12, nwel=5,
                                        nper=1, nstp=1)
    sim, amat, rhs, x0 = oracle_system(case)
    ims = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings)
    ipc = 2 if relax > 0 else 1
    L.orc_pcu(ims.nja, ims.neq, ipc, relax, amat, ims.ia, ims.ja, ims.apc, ims.iapc,
              ims.japc, ims.iwork, ims.work)
    apc, iapc, japc = _ref_factor(sim, amat, ipc, relax)
    assert np.array_equal(ims.iapc + 1, iapc)
    # JAPC(1:NEQ) are positions, JAPC(NEQ+1:) are columns: both shift by 1
    assert np.array_equal(ims.japc + 1, japc)
    assert eqbits(ims.apc, apc)
    r = np.random.default_rng(3).standard_normal(sim.neq)
    z1, z2 = np.zeros(sim.neq), np.zeros(sim.neq)
    L.orc_ilu0a(ims.nja, ims.neq, ims.apc, ims.iapc, ims.japc, r, z1)
    R.ref_ilu0a(sim.neq, ims.nja, apc, iapc, japc, r, z2)
    assert eqbits(z1, z2)


def test_ilu0_early_exit_quirk():
    """A pivot whose sign flips under MILU0 stops pcilu0 at that row;
    IPCFLAG is reset anyway (ImsLinearBase.f90:1040-1041), so the remaining
    rows keep stale APC values."""
    case = configs.steady_small(nlay=1, nrow=6, ncol=6, nwel=2)
    sim, amat, rhs, x0 = oracle_system(case)
    d = sim.dis
    bad = amat.copy()
    bad[d.ia[14]] = 0.5 * amat[d.ia[14]]    # weakened diagonal: a later MILU0 pivot flips
    ims = O.ImsLinear(d.ia, d.ja, sim.settings)
    ims.apc[:] = 0.25             # stale values from an earlier factor
    stop = L.orc_pcu(ims.nja, ims.neq, 2, 0.97, bad, ims.ia, ims.ja, ims.apc, ims.iapc,
                     ims.japc, ims.iwork, ims.work)
    assert stop == 21
    ia1, ja1 = (d.ia + 1).astype(np.int32), (d.ja + 1).astype(np.int32)
    iapc = np.zeros(sim.neq + 1, np.int32)
    japc = np.zeros(d.nja, np.int32)
    apc = np.full(d.nja, 0.25)
    R.ref_pccrs(sim.neq, d.nja, ia1, ja1, iapc, japc)
    R.ref_pcu(sim.neq, d.nja, 2, 0.97, bad, ia1, ja1, apc, iapc, japc)
    assert eqbits(ims.apc, apc)
    assert ims.apc[sim.neq - 1] == 0.25      # the last row was never reached


def _compare_ims(case, north=0, icnvgopt=0, relax=0.0, kstp=1):
    sim, amat, rhs, x0 = oracle_system(case)
    s = sim.settings
    s.north, s.icnvgopt, s.relax = north, icnvgopt, relax
    d = sim.dis
    ims = O.ImsLinear(d.ia, d.ja, s)
    x1 = x0.copy()
    icn1, it1, l21, _, recs = ims.apply(amat, rhs.copy(), x1, kstp=kstp, history=True)
    x2 = x0.copy()
    itmax = s.iter1
    hdv, hr = np.zeros(itmax), np.zeros(itmax)
    hl1, hl2 = np.zeros(itmax, np.int32), np.zeros(itmax, np.int32)
    icn2, it2, l22 = C.c_int(), C.c_int(), C.c_double()
    R.ref_ims_ap(sim.neq, d.nja, (d.ia + 1).astype(np.int32), (d.ja + 1).astype(np.int32),
                 amat, rhs.copy(), x2, s.ilinmeth, s.ipc, icnvgopt, north, itmax,
                 s.dvclose, s.rclose, relax, kstp, C.byref(icn2), C.byref(it2),
                 C.byref(l22), hdv, hr, hl1, hl2)
    assert (icn1, it1) == (icn2.value, it2.value)
    assert eqbits(l21, l22.value)
    assert eqbits(x1, x2)
    assert eqbits([r[0] for r in recs], hdv[:it1])
    assert eqbits([r[1] for r in recs], hr[:it1])
    assert np.array_equal(np.array([r[4] for r in recs]) + 1, hl1[:it1])
    assert np.array_equal(np.array([r[5] for r in recs]) + 1, hl2[:it1])
    return it1


def test_cg_trajectory_bit_exact():
    assert _compare_ims(configs.steady_small()) > 100


@pytest.mark.parametrize("north,icnvgopt,relax", [(2, 0, 0.0), (0, 1, 0.0), (0, 2, 0.0),
                                                  (0, 3, 0.0), (0, 4, 0.0), (3, 0, 0.97)])
def test_cg_options_bit_exact(north, icnvgopt, relax):
    case = configs.steady_small(nlay=3, nrow=30, ncol=24, nwel=8)
    _compare_ims(case, north=north, icnvgopt=icnvgopt, relax=relax)


@pytest.mark.parametrize("kstp", [1, 2])
def test_bicgstab_trajectory_bit_exact(kstp):
    case = configs.transient_newton(seed=8, nlay=3, nrow=24, ncol=20, nwel=8, nper=1,
                                    nstp=2)
    _compare_ims(case, kstp=kstp)


@pytest.mark.parametrize("iscl", [1, 2])
def test_ims_scale_bit_exact(iscl):
    """ims_base_scale (ImsLinearBase.f90:619-754): the scaled A, x, b and both
    scaling vectors, then the unscaled ones, bit for bit; with iscl 1 the
    scaled diagonal is +-1 and with iscl 2 every column has unit L2 norm
    (within rounding)."""
    case = configs.steady_small(nlay=3, nrow=14, ncol=17, nwel=5)
    sim, amat, rhs, x0 = oracle_system(case)
    d = sim.dis
    ia1, ja1 = (d.ia + 1).astype(np.int32), (d.ja + 1).astype(np.int32)
    ia, ja = d.ia.astype(np.int32), d.ja.astype(np.int32)
    n = sim.neq
    a1, x1, b1 = amat.copy(), x0.copy(), rhs.copy()
    a2, x2, b2 = amat.copy(), x0.copy(), rhs.copy()
    s1, t1, s2, t2 = np.ones(n), np.ones(n), np.ones(n), np.ones(n)
    O.lib().orc_ims_scale(0, iscl, n, ia, ja, a1, x1, b1, s1, t1)
    R.ref_ims_scale(0, iscl, n, d.nja, ia1, ja1, a2, x2, b2, s2, t2)
    for u, v in ((a1, a2), (x1, x2), (b1, b2), (s1, s2), (t1, t2)):
        assert eqbits(u, v)
    assert not eqbits(a1, amat)
    if iscl == 1:
        assert np.allclose(np.abs(a1[d.ia[:-1]]), 1.0, rtol=1e-14, atol=0)
    else:
        col = np.zeros(n)
        np.add.at(col, d.ja, a1 * a1)
        assert np.allclose(col, 1.0, rtol=1e-13, atol=0)
    O.lib().orc_ims_scale(1, iscl, n, ia, ja, a1, x1, b1, s1, t1)
    R.ref_ims_scale(1, iscl, n, d.nja, ia1, ja1, a2, x2, b2, s2, t2)
    for u, v in ((a1, a2), (x1, x2), (b1, b2)):
        assert eqbits(u, v)
    assert np.allclose(a1, amat, rtol=1e-14, atol=0)


@pytest.mark.parametrize("iscl,ilinmeth", [(1, 1), (2, 1), (1, 2), (2, 2)])
def test_ims_ap_scaled_bit_exact(iscl, ilinmeth):
    """imslinear_ap with SCALING_METHOD (ImsLinear.f90:646-747): the whole
    scaled CG / BiCGSTAB trajectory (BiCGSTAB's dx and residual maxima in
    unscaled units), the heads and the unscaled A and b, bit for bit."""
    case = configs.steady_small(nlay=3, nrow=30, ncol=24, nwel=8)
    sim, amat, rhs, x0 = oracle_system(case)
    s = sim.settings
    s.iscl, s.ilinmeth = iscl, ilinmeth
    d = sim.dis
    ims = O.ImsLinear(d.ia, d.ja, s)
    a1, b1, x1 = amat.copy(), rhs.copy(), x0.copy()
    icn1, it1, l21, _, recs = ims.apply(a1, b1, x1, history=True)
    itmax = s.iter1
    hdv, hr = np.zeros(itmax), np.zeros(itmax)
    hl1, hl2 = np.zeros(itmax, np.int32), np.zeros(itmax, np.int32)
    icn2, it2, l22 = C.c_int(), C.c_int(), C.c_double()
    a2, b2, x2 = amat.copy(), rhs.copy(), x0.copy()
    R.ref_ims_ap_scl(sim.neq, d.nja, (d.ia + 1).astype(np.int32),
                     (d.ja + 1).astype(np.int32), a2, b2, x2, ilinmeth, s.ipc, 0, 0,
                     itmax, s.dvclose, s.rclose, 0.0, 1, iscl, C.byref(icn2),
                     C.byref(it2), C.byref(l22), hdv, hr, hl1, hl2)
    assert it1 > 10
    assert icn1 == 1 or iscl == 2      # L2-norm scaling may not close here
    assert (icn1, it1) == (icn2.value, it2.value)
    assert eqbits(l21, l22.value)
    for u, v in ((x1, x2), (a1, a2), (b1, b2)):
        assert eqbits(u, v)
    assert eqbits([r[0] for r in recs], hdv[:it1])
    assert eqbits([r[1] for r in recs], hr[:it1])
    assert np.array_equal(np.array([r[4] for r in recs]) + 1, hl1[:it1])
    assert np.array_equal(np.array([r[5] for r in recs]) + 1, hl2[:it1])
    if not icn1:
        return
    # the scaled solve reaches the unscaled one's heads
    ims0 = O.ImsLinear(d.ia, d.ja, O.ImsSettings(s.dvclose, s.rclose, ilinmeth=ilinmeth))
    x3 = x0.copy()
    ims0.apply(amat, rhs, x3)
    assert np.max(np.abs(x3 - x1)) < 10 * s.dvclose
