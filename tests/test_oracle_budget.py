"""Pins the oracle's budget (gwf_cq + gwf_bd: face flows, storage and
package rates, the CHD rate and the model budget terms) with closed forms
and with the balance identity the formulate fixes:

* 1-D flow between two CHD cells: every face carries K A dh / L, CHD in =
  CHD out = that flow, percent discrepancy 0;
* test_gwf_newton01.py's model: recharge 1 on nine unit cells, so RCH in =
  CHD out = 9;
* for a model whose matrix does not depend on the heads (confined, no
  drains), each non-CHD cell's flowja diagonal (net inflow) equals
  (A x - b) of the system formulated at the same heads, the storage rate
  of a confined cell is -Ss thk area (h - hold) / delt, and a steady model
  has no storage term;
* flowja is antisymmetric bit for bit.
"""
import dataclasses

import numpy as np

from modflow6_amd import configs
from oracle import oracle as O
from tests.cases import line_case, newton01_case


def _antisymmetric(sim, fj):
    d = sim.dis
    off = np.ones(d.nja, bool)
    off[d.ia[:-1]] = False
    return np.array_equal(fj[off].view(np.int64), (-fj[d.isym[off]]).view(np.int64))


def test_budget_line_closed_form():
    case = line_case()
    sim = O.GwfSimulation(case)
    sim.run()
    b = sim.budget()
    q = 3.0 * (10.0 * 10.0) / 10.0 * 9.0 / (case.ncol - 1)
    rin, rout = b["terms"]["CHD"]
    assert abs(rin - q) <= 1e-9 * q and abs(rout - q) <= 1e-9 * q
    assert abs(b["pct_discrepancy"]) < 1e-8
    fj = b["flowja"]
    d = sim.dis
    assert _antisymmetric(sim, fj)
    # every face between columns j and j+1 carries q towards j
    for n in range(case.ncol - 1):
        pos = [p for p in range(d.ia[n] + 1, d.ia[n + 1]) if d.ja[p] == n + 1][0]
        assert abs(fj[pos] - q) <= 1e-9 * q
    # interior cells balance; CHD cells' diagonal holds their flow residual 0
    assert np.max(np.abs(fj[d.ia[:-1]])) < 1e-9 * q


def test_budget_newton01_recharge_leaves_through_chd():
    sim = O.GwfSimulation(newton01_case())
    sim.run()
    b = sim.budget()
    assert b["terms"]["RCH"] == (9.0, 0.0)
    rin, rout = b["terms"]["CHD"]
    assert abs(rout - rin - 9.0) < 1e-6
    assert abs(b["pct_discrepancy"]) < 1e-5
    assert _antisymmetric(sim, b["flowja"])


def _confined_transient():
    case = configs.transient_newton(nlay=3, nrow=12, ncol=14, nwel=6, nper=1, nstp=1)
    nodes = case.nodes
    case = dataclasses.replace(
        case, newton=False, newton_ur=False, icelltype=np.zeros(nodes, np.int32),
        iconvert=np.zeros(nodes, np.int32),
        packages=[p for p in case.packages if p.kind != "DRN"])
    return case


def test_budget_diagonal_is_the_system_residual():
    case = _confined_transient()
    sim = O.GwfSimulation(case)
    sim.kper, sim.kstp = 1, 1
    sim.advance(case.periods[0][0], False)
    sim.solve_outer(1)
    sim.solve_outer(2)
    b = sim.budget()
    fj = b["flowja"]
    d = sim.dis
    sim.formulate(3)
    rows = [slice(d.ia[n], d.ia[n + 1]) for n in range(sim.neq)]
    ax = np.array([np.dot(sim.amat[r], sim.x[d.ja[r]]) for r in rows])
    resid = ax - sim.rhs
    # the two sums differ only by rounding: bound by 64 eps times the sum of
    # the magnitudes of each row's terms
    mag = np.array([np.sum(np.abs(sim.amat[r] * sim.x[d.ja[r]])) for r in rows])
    mag = mag + np.abs(sim.rhs)
    assert np.all(np.abs(fj[d.ia[:-1]] - resid) <= 64 * np.finfo(float).eps * mag)
    # confined storage rate: -Ss thk area (h - hold) / delt
    thk = d.top - d.bot
    exp = -case.ss * thk * d.area * (sim.x - sim.xold) / sim.delt
    assert np.allclose(b["strgss"], exp, rtol=1e-12, atol=1e-12 * np.max(np.abs(exp)))
    assert "STO-SY" not in b["terms"]
    ins = sum(v[0] for v in b["terms"].values())
    outs = sum(v[1] for v in b["terms"].values())
    assert abs(ins - outs) <= 1e-6 * max(ins, outs)
    assert _antisymmetric(sim, fj)


def test_budget_steady_has_no_storage_flow():
    case = configs.steady_small(nlay=2, nrow=10, ncol=12, nwel=3)
    sim = O.GwfSimulation(case)
    sim.run()
    b = sim.budget()
    assert "STO-SS" not in b["terms"]
    rch = b["terms"]["RCH"][0]
    wel = b["terms"]["WEL"][1]
    chd_in, chd_out = b["terms"]["CHD"]
    assert abs((rch + chd_in) - (wel + chd_out)) <= 1e-6 * (rch + chd_in)
