"""The oracle's horizontal flow barriers (the HFB package, gwf-hfb.f90).

condsat_modify and condsat_reset (gwf-hfb.f90:770-832), which fold a barrier
into the face's saturated conductance under NEWTON or between confined
cells, are private to the reference's type.  Their arithmetic is pinned by:

* the reference's public hfb_fc (:211-348), which applies the same series
  combination to the matrix between convertible cells without NEWTON: with
  every head above its cell top (so no thickness is limited by the head)
  and the face's npf conductance in amat, its per-barrier conductances
  equal the oracle's condsat after each barrier, bit for bit, for positive,
  zero and negative characteristics and for two barriers on one face;
* a closed form: a confined line of cells between two constant heads with a
  barrier on one face, where the discharge is the head difference over the
  sum of the face resistances, 1 / (1/C + 1/(hydchr w b)) for the barrier's
  face;
* condsat_reset in barrier order restoring the faces (and, with two
  barriers on a face, the value between them, as the reference does)."""
import numpy as np
import pytest

from modflow6_amd import configs
from modflow6_amd.configs import GwfCase, ImsOptions, Package
from oracle import oracle as O

L = O.lib()
R = O.ref_gwf_lib()


def eqbits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.int64),
                          np.asarray(b, dtype=np.float64).view(np.int64))


def one(a):
    return np.ascontiguousarray(np.asarray(a) + 1, dtype=np.int32)


def _barriers(d, rng, nb):
    """nb barriers on random horizontal faces (either orientation), then the
    first face again in the other orientation, with positive, zero and
    negative characteristics: (noden, nodem, ja position of nodem in row
    noden, hydchr)."""
    noden, nodem, idx = [], [], []
    while len(noden) < nb:
        a = int(rng.integers(d.nodes))
        pos = [p for p in range(int(d.ia[a]) + 1, int(d.ia[a + 1])) if d.ihc[d.jas[p]] != 0]
        if not pos:
            continue
        p = pos[int(rng.integers(len(pos)))]
        noden.append(a)
        nodem.append(int(d.ja[p]))
        idx.append(p)
    a, m = nodem[0], noden[0]
    noden.append(a)
    nodem.append(m)
    idx.append(next(p for p in range(int(d.ia[a]) + 1, int(d.ia[a + 1])) if d.ja[p] == m))
    hy = rng.uniform(1e-4, 5e-2, len(noden))
    hy[1] = 0.0
    hy[2] = -0.35
    return (np.array(noden, dtype=np.int32), np.array(nodem, dtype=np.int32),
            np.array(idx, dtype=np.int32), hy)


@pytest.mark.skipif(R is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [3, 4])
def test_hfb_combination_matches_reference_hfb_fc(seed):
    rng = np.random.default_rng(seed)
    case = configs.steady_small(seed=seed, nlay=2, nrow=6, ncol=7, nwel=2)
    d = O.dis_connections(case.nlay, case.nrow, case.ncol, case.delr, case.delc,
                          case.top2d, case.botm)
    n = d.nodes
    k11 = np.ascontiguousarray(case.k11, dtype=np.float64)
    condsat = np.zeros(d.hwva.size)
    L.orc_calc_condsat(n, d.ia, d.ja, d.jas, d.ihc, d.cl1, d.cl2, d.hwva, d.top, d.bot,
                       k11, k11, d.ncol, 0.1 * k11, 0, np.ones(n), np.ascontiguousarray(d.top),
                       condsat)
    noden, nodem, idx, hy = _barriers(d, rng, 6)
    # the matrix hfb_fc starts from: npf conductances off the diagonal
    rows = np.repeat(np.arange(n), np.diff(d.ia))
    amat = rng.standard_normal(d.nja)
    off = d.ja != rows
    amat[off] = condsat[d.jas[off]]
    R.ref_gwf_dis(n, d.nja, d.hwva.size, one(d.ia), one(d.ja), one(d.jas), d.ihc,
                  one(d.isym), d.hwva, d.cl1, d.cl2, d.top, d.bot, d.area)
    icell = np.ones(n, dtype=np.int32)
    hnew = np.ascontiguousarray(d.top + 1.0)
    a_ref = amat.copy()
    condsav = np.zeros(noden.size)
    R.ref_hfb_fc(n, d.nja, noden.size, 0, one(noden), one(nodem), one(idx), hy, icell,
                 np.ones(n, dtype=np.int32), hnew, a_ref, condsav)
    # oracle, one barrier at a time so each barrier's result can be read
    cs = condsat.copy()
    got = np.zeros(noden.size)
    for b in range(noden.size):
        sav = np.zeros(1)
        L.orc_hfb_condsat_modify(1, noden[b:b + 1], nodem[b:b + 1], idx[b:b + 1],
                                 hy[b:b + 1], 1, icell, d.jas, d.ihc, d.hwva, d.top, d.bot,
                                 cs, sav)
        got[b] = cs[d.jas[idx[b]]]
    assert eqbits(got, condsav)
    for b in range(noden.size):
        assert eqbits(a_ref[idx[b]], cs[d.jas[idx[b]]])
        assert eqbits(a_ref[d.isym[idx[b]]], cs[d.jas[idx[b]]])
    # the whole list at once gives the same faces
    cs2 = condsat.copy()
    sav = np.zeros(noden.size)
    L.orc_hfb_condsat_modify(noden.size, noden, nodem, idx, hy, 1, icell, d.jas, d.ihc,
                             d.hwva, d.top, d.bot, cs2, sav)
    assert eqbits(cs2, cs)


def _line(hydchr, face=3, ncol=8, k=3.0):
    nodes = ncol
    return GwfCase(
        name="line_hfb", nlay=1, nrow=1, ncol=ncol, delr=np.full(ncol, 10.0),
        delc=np.full(1, 20.0), top2d=np.zeros((1, ncol)),
        botm=np.full((1, 1, ncol), -5.0), k11=np.full(nodes, k),
        k33=np.full(nodes, k), icelltype=np.zeros(nodes, np.int32),
        strt=np.zeros(nodes),
        packages=[Package("CHD", np.array([0, ncol - 1]), np.array([1.0, 10.0]))],
        ims=ImsOptions(inner_dvclose=1e-13, inner_rclose=1e-13, outer_dvclose=1e-13),
        periods=[(1.0, 1, True)],
        hfb=(np.array([face]), np.array([face + 1]), np.array([hydchr])))


@pytest.mark.parametrize("hydchr", [0.05, 1e-3, -0.25])
def test_hfb_line_closed_form(hydchr):
    """Confined line, K = 3, cells 10 m long, 20 m wide, 5 m thick: each face
    conducts C = K w b / L = 30; the barrier's face 1 / (1/30 + 1/(hydchr w
    b)) for hydchr > 0, or -hydchr * 30 as a multiplier."""
    case = _line(hydchr)
    sim = O.GwfSimulation(case)
    sim.kper = 1
    conv, _ = sim.run_timestep(1.0, True)
    assert conv
    c = 3.0 * 20.0 * 5.0 / 10.0
    cb = 1.0 / (1.0 / c + 1.0 / (hydchr * 20.0 * 5.0)) if hydchr > 0 else -hydchr * c
    res = np.full(case.ncol - 1, 1.0 / c)
    res[3] = 1.0 / cb
    q = (10.0 - 1.0) / res.sum()
    want = 1.0 + q * np.concatenate([[0.0], np.cumsum(res)])
    assert np.allclose(sim.x, want, rtol=0, atol=1e-9 * 10.0)


def test_hfb_reset_restores_faces():
    """hfb_rp with a new (empty) set restores the faces through the saved
    values in barrier order: exactly, except that a face with two barriers
    keeps the value between them (the second barrier saved it last)."""
    case = configs.steady_small(nlay=2, nrow=6, ncol=7, nwel=2)
    sim = O.GwfSimulation(case)
    orig = sim.condsat.copy()
    d = sim.dis
    noden = np.array([3, 10, 3])
    nodem = np.array([4, 17, 4])
    sim.set_hfb(noden, nodem, np.array([0.01, 0.02, 0.03]))
    f1 = d.jas[sim.hfb[1][0]]
    f2 = d.jas[sim.hfb[1][1]]
    after_first = orig[f1] * (0.01 * d.hwva[f1] * 10.0) / (orig[f1] + 0.01 * d.hwva[f1] * 10.0)
    assert sim.condsat[f2] != orig[f2]
    sim.set_hfb(np.array([], dtype=int), np.array([], dtype=int), np.array([]))
    assert sim.condsat[f2] == orig[f2]
    mask = np.ones(orig.size, dtype=bool)
    mask[f1] = False
    assert eqbits(sim.condsat[mask], orig[mask])
    assert np.isclose(sim.condsat[f1], after_first, rtol=1e-15)


def test_hfb_rejects_unconnected_and_vertical_pairs():
    case = configs.steady_small(nlay=2, nrow=6, ncol=7, nwel=2)
    sim = O.GwfSimulation(case)
    with pytest.raises(ValueError, match="unconnected"):
        sim.set_hfb(np.array([0]), np.array([2]), np.array([0.1]))
    with pytest.raises(ValueError, match="not horizontally"):
        sim.set_hfb(np.array([0]), np.array([42]), np.array([0.1]))


@pytest.mark.skipif(R is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [5, 6])
def test_hfb_fc_cq_convertible_without_newton(seed):
    """hfb_fc and hfb_cq between convertible cells without NEWTON
    (gwf-hfb.f90:211-450) against the reference's, with heads below and
    above the cell tops, inactive and constant-head cells, mixed cell types
    and two barriers on one face."""
    rng = np.random.default_rng(seed)
    case = configs.steady_small(seed=seed, nlay=2, nrow=6, ncol=7, nwel=2)
    d = O.dis_connections(case.nlay, case.nrow, case.ncol, case.delr, case.delc,
                          case.top2d, case.botm)
    n = d.nodes
    R.ref_gwf_dis(n, d.nja, d.hwva.size, one(d.ia), one(d.ja), one(d.jas), d.ihc,
                  one(d.isym), d.hwva, d.cl1, d.cl2, d.top, d.bot, d.area)
    noden, nodem, idx, hy = _barriers(d, rng, 10)
    icell = np.ascontiguousarray(rng.integers(0, 2, n), dtype=np.int32)
    ib = np.ones(n, dtype=np.int32)
    ib[rng.choice(n, 4, replace=False)] = -1
    ib[rng.choice(n, 3, replace=False)] = 0
    hnew = np.ascontiguousarray(d.bot + rng.uniform(0.1, 1.3, n) * (d.top - d.bot))
    amat = rng.uniform(0.5, 50.0, d.nja)
    a1, a2 = amat.copy(), amat.copy()
    s1, s2 = np.zeros(noden.size), np.zeros(noden.size)
    L.orc_hfb_fc(noden.size, noden, nodem, idx, hy, ib, icell, d.ia, d.isym, d.jas, d.hwva,
                 d.top, d.bot, hnew, a1, s1)
    R.ref_hfb_fc(n, d.nja, noden.size, 0, one(noden), one(nodem), one(idx), hy, icell, ib,
                 hnew, a2, s2)
    assert eqbits(a1, a2) and eqbits(s1, s2)
    assert not eqbits(a1, amat)
    f1 = rng.standard_normal(d.nja)
    f2 = f1.copy()
    L.orc_hfb_cq(noden.size, noden, nodem, idx, d.isym, ib, icell, hnew, s1, f1)
    R.ref_hfb_cq(n, d.nja, noden.size, one(noden), one(nodem), one(idx), icell, ib, hnew,
                 s1, f2)
    assert eqbits(f1, f2)
