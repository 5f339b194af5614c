"""Pins the C oracle's GWF assembly and outer-loop pieces to the reference's
own Fortran, bit for bit.

oracle/_ref/libmf6refgwf.so is the module closure of gwf.f90 and
NumericalSolution.f90 compiled in place from /root/reference/src by
oracle/Makefile; oracle/ref_gwf_shim.f90 allocates the reference's derived
types (DisBaseType, GwfNpfType, GwfStoType, DrnType, NumericalSolutionType,
GwfModelType, SparseMatrixType, SeqVectorType), fills their members and
calls the type-bound procedures:

* npf_cf / npf_fc / npf_fn (gwf-npf.f90:444, :474, :578), npf_nur (:705)
* npf_cq (gwf-npf.f90:745)
* sto_fc / sto_fn / sto_cq (gwf-sto.f90:226, :353, :447), with STO's
  satomega as source_options sets it (1e-6 under NEWTON, :826-829)
* drn_cf + drn_fc + drn_fn + bnd_cq (gwf-drn.f90:340, :377, :418,
  BoundaryPackage.f90:545), with and without a depth auxiliary
* gnc_fc (explicit) / gnc_fn / gnc_cq of one model (GhostNode.f90:280,
  :340, :478), the reference's GhostNodeType with positions from its gnc_mc
* wel_cf + wel_fc + wel_fn + bnd_cq (gwf-wel.f90:296, :340, :378), with and
  without AUTO_FLOW_REDUCE (WelType is private to its module: the shim makes
  it with wel_create and fills it through the memory manager)
* sln_underrelax (NumericalSolution.f90:2989): SIMPLE, COOLEY and
  delta-bar-delta over several outer iterations
* sln_backtracking_xupdate (:2783)
* the residual of sln_calc_residual (:2966), its SeqVector norm2, and
  gwf_ptc (gwf.f90:625)

Every term is exercised with convertible and confined cells, inactive and
constant-head cells, heads below the cell bottom, inside the smoothing band
and above the top, and nonzero starting amat / rhs so the accumulation order
is compared too.  Skipped when the library has not been built (it needs the
reference sources)."""
import ctypes as C

import numpy as np
import pytest

from modflow6_amd import configs
from oracle import oracle as O

R = O.ref_gwf_lib()
L = O.lib()
pytestmark = pytest.mark.skipif(R is None, reason="oracle/_ref not built")


def eqbits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.int64),
                          np.asarray(b, dtype=np.float64).view(np.int64))


def one(a):
    return np.ascontiguousarray(np.asarray(a) + 1, dtype=np.int32)


class Grid:
    """A transient Newton model (the config-2 recipe at a small size) with
    randomised state: mixed cell types, some inactive and constant-head
    cells, heads spread across every saturation regime."""

    def __init__(self, seed, nlay=4, nrow=9, ncol=11):
        rng = np.random.default_rng(seed)
        case = configs.transient_newton(seed=seed, nlay=nlay, nrow=nrow,
                                        ncol=ncol, nwel=6, nper=1, nstp=1)
        self.sim = O.GwfSimulation(case)
        s = self.sim
        d = s.dis
        self.d = d
        n = d.nodes
        self.n = n
        self.icelltype = np.ascontiguousarray(rng.integers(0, 2, n), dtype=np.int32)
        ib = np.ones(n, dtype=np.int32)
        ib[rng.choice(n, n // 12, replace=False)] = 0
        ib[rng.choice(n, n // 12, replace=False)] = -1
        self.ibound = ib
        thk = d.top - d.bot
        u = rng.uniform(-0.4, 1.4, n)
        u[rng.choice(n, n // 10, replace=False)] = rng.uniform(-1e-7, 1e-7, n // 10)
        u[rng.choice(n, n // 10, replace=False)] = 1.0 + rng.uniform(-1e-7, 1e-7, n // 10)
        self.hnew = np.ascontiguousarray(d.bot + u * thk)
        self.hold = np.ascontiguousarray(d.bot + rng.uniform(-0.3, 1.3, n) * thk)
        self.k11 = np.ascontiguousarray(s.k11)
        self.k33 = np.ascontiguousarray(0.1 * s.k11)
        self.condsat = np.zeros(d.hwva.size)
        L.orc_calc_condsat(n, d.ia, d.ja, d.jas, d.ihc, d.cl1, d.cl2, d.hwva,
                           d.top, d.bot, self.k11, self.k11, d.ncol, self.k33, 0,
                           np.ones(n), np.ascontiguousarray(d.top), self.condsat)
        self.sat = np.ascontiguousarray(rng.uniform(0.0, 1.0, n))
        self.amat0 = rng.standard_normal(d.nja)
        self.rhs0 = rng.standard_normal(n)
        self.rng = rng
        R.ref_gwf_dis(n, d.nja, d.hwva.size, one(d.ia), one(d.ja), one(d.jas),
                      d.ihc, one(d.isym), d.hwva, d.cl1, d.cl2, d.top, d.bot,
                      d.area)


@pytest.fixture(params=[11, 12, 13])
def grid(request):
    return Grid(request.param)


@pytest.mark.parametrize("inewton,satomega", [(0, 0.0), (1, 1e-6), (1, 1e-3)])
def test_npf_cf(grid, inewton, satomega):
    """Without NEWTON, npf_cf also dries convertible cells whose head is at
    or below the bottom (sgwf_npf_wetdry, gwf-npf.f90:2061); the oracle and
    the device reject that case, so the heads here keep those cells wet."""
    g = grid
    h = g.hnew.copy()
    if inewton == 0:
        d = g.d
        h = np.ascontiguousarray(np.maximum(h, d.bot + 1e-3 * (d.top - d.bot)))
    s1, s2 = g.sat.copy(), g.sat.copy()
    L.orc_npf_cf(g.n, inewton, satomega, g.icelltype, g.ibound, g.d.top,
                 g.d.bot, h, s1)
    R.ref_npf_cf(g.n, inewton, satomega, g.icelltype, g.ibound, h, s2)
    assert eqbits(s1, s2)
    assert not eqbits(s1, g.sat)


@pytest.mark.parametrize("inewton,ivarcv,idewatcv",
                         [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)])
def test_npf_fc(grid, inewton, ivarcv, idewatcv):
    g, d = grid, grid.d
    sat = g.sat.copy()
    L.orc_npf_cf(g.n, inewton, 1e-6, g.icelltype, g.ibound, d.top, d.bot,
                 g.hnew, sat)
    a1, a2 = g.amat0.copy(), g.amat0.copy()
    r2 = g.rhs0.copy()
    L.orc_npf_fc(g.n, d.ia, d.ja, d.jas, d.isym, d.ihc, d.cl1, d.cl2, d.hwva,
                 g.condsat, g.ibound, g.icelltype, inewton, ivarcv, idewatcv, 0,
                 g.k11, g.k11, d.ncol, g.k33, d.top, d.bot, sat, g.hnew, a1)
    R.ref_npf_fill(0, g.n, d.nja, 1, inewton, ivarcv, idewatcv, 0, 1e-6,
                   g.icelltype, g.ibound, g.k11, g.k33, g.condsat, sat, g.hnew,
                   a2, r2)
    assert eqbits(a1, a2)
    assert eqbits(r2, g.rhs0)          # npf_fc leaves rhs alone (no perched)
    assert not eqbits(a1, g.amat0)


@pytest.mark.parametrize("ivarcv,satomega", [(0, 1e-6), (1, 1e-6), (0, 1e-2)])
def test_npf_fn(grid, ivarcv, satomega):
    g, d = grid, grid.d
    a1, a2 = g.amat0.copy(), g.amat0.copy()
    r1, r2 = g.rhs0.copy(), g.rhs0.copy()
    L.orc_npf_fn(g.n, d.ia, d.ja, d.jas, d.isym, d.ihc, g.condsat, g.ibound,
                 g.icelltype, ivarcv, d.top, d.bot, satomega, g.hnew, a1, r1)
    R.ref_npf_fill(1, g.n, d.nja, 2, 1, ivarcv, 1, 0, satomega, g.icelltype,
                   g.ibound, g.k11, g.k33, g.condsat, g.sat, g.hnew, a2, r2)
    assert eqbits(a1, a2)
    assert eqbits(r1, r2)
    assert not eqbits(r1, g.rhs0)      # Newton terms were added


@pytest.mark.parametrize("delt,satomega", [(30.0 / 5, 1e-6), (1.0, 0.0), (0.37, 1e-6)])
@pytest.mark.parametrize("istor,iconf,iorig", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                                               (1, 1, 0)])
def test_sto_fc_fn_cq(grid, delt, satomega, istor, iconf, iorig):
    """sto_fc / sto_fn / sto_cq with STORAGECOEFFICIENT, SS_CONFINED_ONLY and
    the original specific storage formulation."""
    g, d = grid, grid.d
    rng = g.rng
    iconvert = np.ascontiguousarray(rng.integers(0, 2, g.n), dtype=np.int32)
    ss = rng.uniform(1e-6, 1e-4, g.n)
    sy = rng.uniform(0.05, 0.3, g.n)
    for iflag in (0, 1):
        a1, a2 = g.amat0.copy(), g.amat0.copy()
        r1, r2 = g.rhs0.copy(), g.rhs0.copy()
        if iflag == 0:
            L.orc_sto_fc(g.n, d.ia, g.ibound, iconvert, d.top, d.bot, d.area,
                         ss, sy, istor, iconf, iorig, delt, satomega, g.hold, g.hnew, a1, r1)
        else:
            L.orc_sto_fn(g.n, d.ia, g.ibound, iconvert, d.top, d.bot, d.area,
                         ss, sy, istor, iconf, iorig, delt, g.hnew, a1, r1)
        R.ref_sto_fill(iflag, g.n, d.nja, 2, delt, 0, istor, iconf, iorig, satomega,
                       iconvert, g.ibound, ss, sy, g.hold, g.hnew, a2, r2)
        assert eqbits(a1, a2), iflag
        assert eqbits(r1, r2), iflag
        assert not eqbits(r1, g.rhs0)
    f1, f2 = g.amat0.copy(), g.amat0.copy()
    s1, y1, s2, y2 = (np.zeros(g.n) for _ in range(4))
    L.orc_sto_cq(g.n, d.ia, g.ibound, iconvert, d.top, d.bot, d.area, ss, sy,
                 istor, iconf, iorig, delt, satomega, g.hold, g.hnew, f1, s1, y1)
    R.ref_sto_cq(g.n, d.nja, delt, 0, istor, iconf, iorig, satomega, iconvert, g.ibound,
                 ss, sy,
                 g.hold, g.hnew, f2, s2, y2)
    assert eqbits(f1, f2) and eqbits(s1, s2) and eqbits(y1, y2)
    assert np.count_nonzero(y1) > 0


def test_sto_satomega_matters(grid):
    """A head within the smoothing band at a cell top: satomega 0 and 1e-6
    give different storage terms, so the test above can tell them apart."""
    g, d = grid, grid.d
    h = g.hnew.copy()
    h[:] = d.top - 1e-7 * (d.top - d.bot)
    ic = np.ones(g.n, dtype=np.int32)
    out = []
    for om in (0.0, 1e-6):
        a, r = g.amat0.copy(), g.rhs0.copy()
        R.ref_sto_fill(0, g.n, d.nja, 1, 1.0, 0, 0, 0, 0, om, ic, g.ibound,
                       np.full(g.n, 1e-5), np.full(g.n, 0.2), g.hold, h, a, r)
        out.append(r)
    assert not eqbits(out[0], out[1])


def test_sto_steady_is_noop(grid):
    g, d = grid, grid.d
    a2, r2 = g.amat0.copy(), g.rhs0.copy()
    z = np.ones(g.n, dtype=np.int32)
    R.ref_sto_fill(0, g.n, d.nja, 1, 1.0, 1, 0, 0, 0, 0.0, z, g.ibound,
                   np.full(g.n, 1e-5), np.full(g.n, 0.2), g.hold, g.hnew, a2, r2)
    assert eqbits(a2, g.amat0) and eqbits(r2, g.rhs0)


@pytest.mark.parametrize("idepth,icubic", [(0, 0), (1, 0), (1, 1)])
def test_drn_cf_fc_fn_cq(grid, idepth, icubic):
    """drn_cf + drn_fc + drn_fn + bnd_cq: a step at the drain elevation, or
    with a depth auxiliary (AUXDEPTHNAME; positive, negative and zero
    depths) the linear (sQuadraticSaturation, eps = 0) or cubic
    (sQSaturation, DEV_CUBIC_SCALING) scaling and the cubic Newton term."""
    g, d = grid, grid.d
    rng = g.rng
    nb = g.n // 3
    nodes = np.ascontiguousarray(rng.choice(g.n, nb, replace=False), dtype=np.int32)
    elev = g.hnew[nodes] + rng.uniform(-2.0, 2.0, nb)
    elev[: nb // 6] = g.hnew[nodes[: nb // 6]]        # xnew == drain elevation
    cond = rng.uniform(1.0, 500.0, nb)
    depth = rng.uniform(-3.0, 3.0, nb)
    depth[nb // 6: nb // 4] = 0.0
    h1, b1 = np.zeros(nb), np.zeros(nb)
    L.orc_drn_cf(nb, nodes, elev, cond, idepth, depth, icubic, g.ibound, g.hnew, h1, b1)
    a1, r1 = g.amat0.copy(), g.rhs0.copy()
    L.orc_bnd_fc(nb, nodes, h1, b1, d.ia, a1, r1)
    L.orc_drn_fn(nb, nodes, elev, cond, idepth, depth, g.ibound, g.hnew, d.ia, a1, r1)
    f1, sv1 = g.amat0.copy(), np.zeros(nb)
    L.orc_bnd_cq(nb, nodes, h1, b1, g.ibound, g.hnew, d.ia, f1, sv1)
    h2, b2 = np.zeros(nb), np.zeros(nb)
    a2, r2 = g.amat0.copy(), g.rhs0.copy()
    f2, sv2 = g.amat0.copy(), np.zeros(nb)
    R.ref_drn(g.n, d.nja, nb, one(nodes), elev, cond, idepth, depth, icubic,
              g.ibound, g.hnew, h2, b2, a2, r2, f2, sv2)
    assert eqbits(h1, h2) and eqbits(b1, b2)
    assert eqbits(a1, a2) and eqbits(r1, r2)
    assert eqbits(f1, f2) and eqbits(sv1, sv2)
    assert np.count_nonzero(h1) > 0 and np.count_nonzero(h1) < nb
    if idepth:
        # some drains inside their scaling interval: fractional factors
        frac = -h1 / np.where(g.ibound[nodes] > 0, cond, 1.0)
        assert np.count_nonzero((frac > 0.0) & (frac < 1.0)) > 0


@pytest.mark.parametrize("iflowred,flowred", [(0, 0.0), (1, 0.05), (1, 0.4), (1, 1.0)])
def test_wel_cf_fc_fn_cq(grid, iflowred, flowred):
    """wel_cf + wel_fc + wel_fn + bnd_cq (gwf-wel.f90:296-424) with and
    without AUTO_FLOW_REDUCE: withdrawals in convertible cells scaled by
    sQSaturation over the bottom flowred of the cell, injections and
    confined cells untouched, the Newton term on the diagonal and rhs."""
    g, d = grid, grid.d
    rng = g.rng
    nb = g.n // 3
    nodes = np.ascontiguousarray(rng.choice(g.n, nb, replace=False), dtype=np.int32)
    q = rng.uniform(-400.0, 150.0, nb)
    q[: nb // 10] = 0.0
    # a fifth of the wells with heads inside the reduction interval
    h = g.hnew.copy()
    band = nodes[nb // 10: nb // 10 + nb // 5]
    fr = flowred if iflowred else 0.3
    h[band] = d.bot[band] + rng.uniform(0.05, 0.95, band.size) * fr * (d.top[band] - d.bot[band])
    h1, b1 = np.zeros(nb), np.zeros(nb)
    L.orc_wel_cf(nb, nodes, q, g.ibound, iflowred, flowred, g.icelltype, d.top, d.bot,
                 h, h1, b1)
    a1, r1 = g.amat0.copy(), g.rhs0.copy()
    L.orc_bnd_fc(nb, nodes, h1, b1, d.ia, a1, r1)
    L.orc_wel_fn(nb, nodes, q, b1, g.ibound, iflowred, flowred, g.icelltype, d.top,
                 d.bot, h, d.ia, a1, r1)
    f1, sv1 = g.amat0.copy(), np.zeros(nb)
    L.orc_bnd_cq(nb, nodes, h1, b1, g.ibound, h, d.ia, f1, sv1)
    h2, b2 = np.zeros(nb), np.zeros(nb)
    a2, r2 = g.amat0.copy(), g.rhs0.copy()
    f2, sv2 = g.amat0.copy(), np.zeros(nb)
    R.ref_wel(g.n, d.nja, nb, one(nodes), q, iflowred, flowred, g.icelltype, g.ibound,
              h, h2, b2, a2, r2, f2, sv2)
    assert eqbits(h1, h2) and eqbits(b1, b2)
    assert eqbits(a1, a2) and eqbits(r1, r2)
    assert eqbits(f1, f2) and eqbits(sv1, sv2)
    act = g.ibound[nodes] > 0
    if iflowred:
        # reduced withdrawals: some between zero and the full rate, and the
        # Newton term changed the matrix
        red = act & (q < 0) & (g.icelltype[nodes] != 0)
        frac = -b1[red] / q[red]
        assert np.count_nonzero((frac > 0.0) & (frac < 1.0)) > 0
        assert not eqbits(a1, g.amat0)
    else:
        assert eqbits(b1[act], -q[act]) and eqbits(a1, g.amat0)


def _gncs(g, rng, nexg, numjs=3):
    """nexg ghost-node corrections on random connections (either direction,
    some vertical), contributing cells among the neighbours of n and
    elsewhere, unused slots (-1) and zero weights."""
    d = g.d
    noden, nodem, js, al = [], [], [], []
    while len(noden) < nexg:
        n = int(rng.integers(g.n))
        row = d.ja[d.ia[n] + 1:d.ia[n + 1]]
        if row.size == 0:
            continue
        m = int(row[rng.integers(row.size)])
        cand = np.concatenate([row, rng.integers(0, g.n, 2)])
        j = cand[rng.integers(cand.size, size=numjs)].astype(np.int32)
        j[rng.random(numjs) < 0.2] = -1
        a = rng.uniform(-0.3, 0.6, numjs)
        a[rng.random(numjs) < 0.15] = 0.0
        noden.append(n)
        nodem.append(m)
        js.append(j)
        al.append(a)
    return (np.array(noden, dtype=np.int32), np.array(nodem, dtype=np.int32),
            np.ascontiguousarray(js, dtype=np.int32), np.ascontiguousarray(al))


@pytest.mark.parametrize("inewton", [0, 1])
def test_gnc_fc_fn_cq(grid, inewton):
    """gnc_fc (explicit, with gnc_fmsav's conductances read from the matrix),
    gnc_fn and gnc_cq (GhostNode.f90:251-543) of one model against the
    reference's GhostNodeType (positions from its own gnc_mc)."""
    g, d = grid, grid.d
    rng = g.rng
    nexg, numjs = g.n // 4, 3
    gn, gm, gj, ga = _gncs(g, rng, nexg, numjs)
    ipos = np.zeros(nexg, dtype=np.int32)
    for k in range(nexg):
        row = d.ja[d.ia[gn[k]] + 1:d.ia[gn[k] + 1]]
        ipos[k] = d.ia[gn[k]] + 1 + np.flatnonzero(row == gm[k])[0]
    a1, r1 = g.amat0.copy(), g.rhs0.copy()
    c1 = np.zeros(nexg)
    L.orc_gnc_fc(nexg, numjs, gn, gm, gj, ga, ipos, g.ibound, g.hnew, a1, c1, r1)
    if inewton:
        L.orc_gnc_fn(nexg, numjs, gn, gm, gj, ga, d.ia, ipos, d.isym, d.jas, d.ihc,
                     g.condsat, 0, g.icelltype, d.top, d.bot, g.ibound, g.hnew, a1, r1)
    f1 = g.amat0.copy()
    L.orc_gnc_cq(nexg, numjs, gn, gm, gj, ga, ipos, d.isym, g.ibound, g.hnew, c1, f1)
    a2, r2 = g.amat0.copy(), g.rhs0.copy()
    c2, f2 = np.zeros(nexg), g.amat0.copy()
    R.ref_gnc(g.n, d.nja, nexg, numjs, one(gn), one(gm),
              np.ascontiguousarray(gj + 1, dtype=np.int32), ga, g.ibound, g.hnew,
              inewton, 0, g.icelltype, g.condsat, a2, r2, c2, f2)
    assert eqbits(c1, c2)
    assert eqbits(r1, r2) and eqbits(a1, a2)
    assert eqbits(f1, f2)
    assert not eqbits(r1, g.rhs0) and not eqbits(f1, g.amat0)
    if inewton:
        assert not eqbits(a1, g.amat0)


@pytest.mark.parametrize("inewton", [0, 1])
def test_npf_cq(grid, inewton):
    g, d = grid, grid.d
    sat = g.sat.copy()
    h = g.hnew.copy()
    if inewton == 0:
        h = np.ascontiguousarray(np.maximum(h, d.bot + 1e-3 * (d.top - d.bot)))
    L.orc_npf_cf(g.n, inewton, 1e-6, g.icelltype, g.ibound, d.top, d.bot, h, sat)
    f1, f2 = g.amat0.copy(), g.amat0.copy()
    L.orc_npf_cq(g.n, d.ia, d.ja, d.jas, d.isym, d.ihc, d.cl1, d.cl2, d.hwva,
                 g.condsat, g.ibound, g.icelltype, inewton, 0, 1, 0, g.k11,
                 g.k11, d.ncol, g.k33, d.top, d.bot, sat, h, f1)
    R.ref_npf_cq(g.n, d.nja, inewton, 0, 1, 0, 1e-6, g.icelltype, g.ibound,
                 g.k11, g.k33, g.condsat, sat, h, f2)
    assert eqbits(f1, f2)


def test_npf_nur(grid):
    g = grid
    s = g.sim
    ibotnode = s._ibotnode()
    icelltype = g.icelltype
    x0 = g.hnew.copy()
    xtemp = g.hold.copy()
    dx1 = g.rng.standard_normal(g.n)
    dx2 = dx1.copy()
    x1, x2 = x0.copy(), x0.copy()
    dm1, lm1 = C.c_double(0.0), C.c_int(-1)
    dm2, lm2 = C.c_double(0.0), C.c_int(0)
    i1 = L.orc_npf_nur(g.n, g.ibound, icelltype, ibotnode, g.d.bot, x1, xtemp,
                       dx1, C.byref(dm1), C.byref(lm1))
    i2 = R.ref_npf_nur(g.n, icelltype, g.ibound, one(ibotnode), x2, xtemp, dx2,
                       C.byref(dm2), C.byref(lm2))
    assert i1 == i2 == 1
    assert eqbits(x1, x2) and eqbits(dx1, dx2)
    assert eqbits(dm1.value, dm2.value) and lm1.value + 1 == lm2.value


@pytest.mark.parametrize("nonmeth,gamma,theta,kappa,momentum", [
    (1, 0.7, 1.0, 0.0, 0.0),
    (2, 0.9, 1.0, 0.0, 0.0),
    (2, 0.0, 1.0, 0.0, 0.0),
    (3, 0.2, 0.7, 0.0001, 0.0),
    (3, 0.0, 0.85, 0.05, 0.1),
])
def test_sln_underrelax_sequence(nonmeth, gamma, theta, kappa, momentum):
    """Ten outer iterations with changes that flip sign and shrink, so the
    Cooley factor takes both of its branches and delta-bar-delta both its
    flip-flop and its same-sign updates (and the momentum term after
    iteration 4)."""
    rng = np.random.default_rng(100 + nonmeth)
    n = 400
    active = np.ones(n, dtype=np.int32)
    active[rng.choice(n, 40, replace=False)] = 0
    st = O.UrState(nonmeth, gamma, theta, kappa, momentum, 0.0, 0.0, 0.0)
    dxo1, ws1, hc1, de1 = (np.zeros(n) for _ in range(4))
    R.ref_sln_new(n, nonmeth, gamma, theta, kappa, momentum, 0.0, 0.0)
    scal = np.zeros(3)
    dxo2, ws2, hc2, de2 = (np.zeros(n) for _ in range(4))
    xtemp = rng.uniform(0.0, 50.0, n)
    relax_branches = set()
    for kiter in range(1, 11):
        sign = -1.0 if kiter % 3 == 0 else 1.0
        step = sign * rng.uniform(0.1, 2.0, n) / kiter
        x = xtemp + step
        bigch = float(step[np.argmax(np.abs(step))])
        x1, x2 = x.copy(), x.copy()
        L.orc_underrelax(C.byref(st), kiter, bigch, n, active, x1, xtemp,
                         dxo1, ws1, hc1, de1)
        R.ref_sln_underrelax(kiter, bigch, n, active, x2, xtemp, scal, dxo2,
                             ws2, hc2, de2)
        assert eqbits(x1, x2), kiter
        assert eqbits(dxo1, dxo2), kiter
        if nonmeth == 2:
            assert eqbits([st.bigch, st.bigchold, st.relaxold], scal), kiter
            relax_branches.add(st.relaxold < 1.0)
        if nonmeth == 3:
            assert eqbits(ws1, ws2) and eqbits(hc1, hc2) and eqbits(de1, de2)
        xtemp = x1
    if nonmeth == 2:
        assert relax_branches == {True, False}
    R.ref_sln_free()


@pytest.mark.parametrize("breduc,dvclose", [(0.1, 1e-3), (0.5, 10.0), (0.9, 1e-6)])
def test_backtracking_xupdate(breduc, dvclose):
    rng = np.random.default_rng(7)
    n = 300
    active = np.ones(n, dtype=np.int32)
    active[::17] = 0
    xtemp = rng.uniform(-5.0, 5.0, n)
    x = xtemp + rng.standard_normal(n) * 3.0
    R.ref_sln_new(n, 0, 1.0, 1.0, 0.0, 0.0, breduc, dvclose)
    x1, x2 = x.copy(), x.copy()
    f1 = L.orc_bt_flag(n, active, x1, xtemp, breduc, dvclose)
    if f1:
        L.orc_bt_apply(n, active, x1, xtemp, breduc)
    f2 = R.ref_sln_bt_xupdate(n, active, x2, xtemp)
    assert f1 == f2
    assert eqbits(x1, x2)
    R.ref_sln_free()


@pytest.mark.parametrize("iss,inewton,npf_inewton,zero_resid", [
    (1, 1, 1, False), (1, 0, 1, False), (0, 1, 1, False), (1, 1, 1, True),
    (1, 0, 0, False)])
def test_residual_l2norm_and_gwf_ptc(grid, iss, inewton, npf_inewton, zero_resid):
    g, d = grid, grid.d
    amat = g.amat0.copy()
    x = g.hnew.copy()
    rhs = g.rhs0.copy()
    active = np.ascontiguousarray(np.where(g.ibound > 0, 1, 0), dtype=np.int32)
    if zero_resid:
        y = np.zeros(g.n)
        L.orc_amux(g.n, x, y, amat, d.ja, d.ia)
        rhs = y.copy()
    r1 = np.zeros(g.n)
    l2a = L.orc_sln_l2norm(g.n, d.ia, d.ja, amat, x, rhs, active, r1)
    pf1 = C.c_double(0.0)
    ip1 = L.orc_gwf_ptc(g.n, iss, inewton if inewton else npf_inewton,
                        g.ibound, d.area, d.top, d.bot, r1, 6.0, C.byref(pf1))
    r2 = np.zeros(g.n)
    scal = np.zeros(2)
    ip2 = R.ref_gwf_ptc(g.n, d.nja, amat, x, rhs, active, iss, inewton,
                        npf_inewton, g.ibound, 6.0, r2, scal)
    assert eqbits(r1, r2)
    assert eqbits(l2a, scal[0])
    assert ip1 == ip2
    assert eqbits(pf1.value, scal[1])
    if zero_resid and ip1:
        assert pf1.value == 1.0 / (6.0 * 10.0)
