"""Ghost-node correction (explicit GNC of one model, mf6hip_gwf_set_gnc)
against the oracle's GNC (pinned to the reference's GhostNodeType in
tests/test_oracle_gwf_vs_reference.py): the formulated system bit for bit
(gnc_fc with the conductances gnc_fmsav reads, and gnc_fn under NEWTON),
the budget's face flows bit for bit (gnc_cq), the heads of a solved step
within 10 x dvclose, a new set replacing the last one, and the
rejections."""
import dataclasses

import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _gncs(case, seed, nexg=60, numjs=3):
    """nexg corrections on random connections (all six directions), the
    contributing cells among n's neighbours and elsewhere, unused slots and
    zero weights."""
    rng = np.random.default_rng(seed)
    nrc = case.nrow * case.ncol
    noden, nodem, js, al = [], [], [], []
    while len(noden) < nexg:
        n = int(rng.integers(case.nodes))
        k, r = divmod(n, nrc)
        i, j = divmod(r, case.ncol)
        nb = []
        if k > 0:
            nb.append(n - nrc)
        if k < case.nlay - 1:
            nb.append(n + nrc)
        if i > 0:
            nb.append(n - case.ncol)
        if i < case.nrow - 1:
            nb.append(n + case.ncol)
        if j > 0:
            nb.append(n - 1)
        if j < case.ncol - 1:
            nb.append(n + 1)
        m = nb[int(rng.integers(len(nb)))]
        cand = np.array(nb + [int(rng.integers(case.nodes))])
        jj = cand[rng.integers(cand.size, size=numjs)]
        jj[rng.random(numjs) < 0.2] = -1
        a = rng.uniform(-0.2, 0.5, numjs)
        a[rng.random(numjs) < 0.15] = 0.0
        noden.append(n)
        nodem.append(m)
        js.append(jj)
        al.append(a)
    return (np.array(noden), np.array(nodem), np.array(js, dtype=np.int64),
            np.array(al, dtype=np.float64))


def _confined():
    c = configs.steady_small(nlay=3, nrow=20, ncol=24, nwel=6)
    return dataclasses.replace(c, gnc=_gncs(c, 1))


def _newton():
    c = configs.transient_newton(nlay=3, nrow=24, ncol=30, nwel=8, nper=1, nstp=2)
    return dataclasses.replace(c, gnc=_gncs(c, 2))


CASES = {"confined": _confined, "newton": _newton}


@pytest.mark.parametrize("name", sorted(CASES))
def test_formulate_and_budget_with_gnc_bit_exact(name):
    case = CASES[name]()
    sim = O.GwfSimulation(case)
    perlen, nstp, steady = case.periods[0]
    sim.kper, sim.kstp = 1, 1
    sim.advance(perlen / nstp, steady)
    sim.formulate(1)
    gs = GwfSolution(case)
    try:
        gs.kper, gs.kstp = 1, 1
        gs.prepare_timestep(perlen / nstp, steady)
        amat, rhs = gs.formulate(1, want_system=True)
        db = gs.budget()
    finally:
        gs.destroy()
    plain = O.GwfSimulation(dataclasses.replace(case, gnc=None))
    plain.kper, plain.kstp = 1, 1
    plain.advance(perlen / nstp, steady)
    plain.formulate(1)
    assert not np.array_equal(plain.rhs, sim.rhs)
    bad = np.flatnonzero(amat.view(np.int64) != sim.amat.view(np.int64))
    assert bad.size == 0, (bad[:8], amat[bad[:8]], sim.amat[bad[:8]])
    bad = np.flatnonzero(rhs.view(np.int64) != sim.rhs.view(np.int64))
    assert bad.size == 0, (bad[:8], rhs[bad[:8]], sim.rhs[bad[:8]])
    ob = sim.budget()
    fo, fd = ob["flowja"], db["flowja"]
    bad = np.flatnonzero(fo.view(np.int64) != fd.view(np.int64))
    assert bad.size == 0, (bad[:8], fo[bad[:8]], fd[bad[:8]])


@pytest.mark.parametrize("name", sorted(CASES))
def test_step_with_gnc_heads(name):
    case = CASES[name]()
    sim = O.GwfSimulation(case)
    perlen, nstp, steady = case.periods[0]
    sim.kper = 1
    conv, _ = sim.run_timestep(perlen / nstp, steady)
    assert conv
    gs = GwfSolution(case)
    try:
        gs.kper = 1
        gs.run_timestep(perlen / nstp, steady)
        hd = gs.heads()
    finally:
        gs.destroy()
    tol = 10 * max(case.ims.inner_dvclose, case.ims.outer_dvclose)
    print(name, "max |dh|", np.max(np.abs(hd - sim.x)), "tolerance", tol)
    assert np.max(np.abs(hd - sim.x)) < tol


def test_new_gnc_set_replaces_the_last():
    case = _confined()
    g2 = _gncs(case, 9, nexg=20, numjs=2)
    gs = GwfSolution(case)
    sim = O.GwfSimulation(case)
    try:
        gs.kper, gs.kstp = 1, 1
        gs.prepare_timestep(1.0, True)
        sim.kper, sim.kstp = 1, 1
        sim.advance(1.0, True)
        for new in (g2, None):
            if new is None:
                e = np.array([], dtype=int)
                gs.set_gnc(e, e, np.zeros((0, 1), dtype=int), np.zeros((0, 1)))
                sim.gnc = None
            else:
                gs.set_gnc(*new)
                sim.set_gnc(*new)
            sim.formulate(1)
            amat, rhs = gs.formulate(1, want_system=True)
            assert np.array_equal(amat.view(np.int64), sim.amat.view(np.int64))
            assert np.array_equal(rhs.view(np.int64), sim.rhs.view(np.int64))
    finally:
        gs.destroy()


def test_set_gnc_rejections():
    case = configs.steady_small(nlay=2, nrow=6, ncol=7, nwel=2)
    gs = GwfSolution(case)
    try:
        L = gs.L

        def call(n, m, js, al, implicit=0):
            n = np.ascontiguousarray(np.asarray(n) + 1, dtype=np.int32)
            m = np.ascontiguousarray(np.asarray(m) + 1, dtype=np.int32)
            js = np.ascontiguousarray(np.asarray(js) + 1, dtype=np.int32)
            al = np.ascontiguousarray(al, dtype=np.float64)
            return L.mf6hip_gwf_set_gnc(gs.h, n.size, js.size // max(n.size, 1), _lib.iptr(n),
                                        _lib.iptr(m), _lib.iptr(js), _lib.dptr(al), implicit)
        assert call([0], [1], [7], [0.3], implicit=1) == _lib.MF6HIP_EUNSUPPORTED
        assert "EXPLICIT" in _lib.last_error()
        assert call([0], [2], [7], [0.3]) == _lib.MF6HIP_EINVAL
        assert "not connected" in _lib.last_error()
        assert call([0], [1], [84], [0.3]) == _lib.MF6HIP_EINVAL
        assert "out of range" in _lib.last_error()
        assert call([0], [42], [7], [0.3]) == _lib.MF6HIP_OK     # vertical neighbour
    finally:
        gs.destroy()
