"""WEL AUTO_FLOW_REDUCE and DRN AUXDEPTHNAME / DEV_CUBIC_SCALING on the
device (mf6hip_gwf_set_package_ex) against the oracle's packages (pinned to
the reference's wel / drn in tests/test_oracle_gwf_vs_reference.py): the
formulated system bit for bit with and without NEWTON (reduced withdrawals,
linear and cubic drain scaling and their Newton terms), the budget's face
flows and per-bound rates bit for bit, and the heads of solved time steps
within 10 x dvclose."""
import dataclasses

import numpy as np
import pytest

from modflow6_amd import configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(newton, cubic=False, flowred=0.5, seed=4):
    """A transient model, all cells convertible: its wells with
    AUTO_FLOW_REDUCE and heads inside the reduction interval, its drains
    with depths of both signs (a fifth of them zero) around heads half a
    metre below the top."""
    c = configs.transient_newton(seed=seed, nlay=3, nrow=24, ncol=30, nwel=16, nper=1,
                                 nstp=2)
    rng = np.random.default_rng(seed)
    wel, drn, rch = c.packages
    bot = c.botm.reshape(-1)
    top = np.concatenate([c.top2d.reshape(-1), c.botm[:-1].reshape(-1)])
    strt = c.strt.copy()
    wn = np.asarray(wel.nodes)
    # heads inside the interval [bot, bot + flowred (top - bot)] for most
    # wells, above it for the rest
    fr = configs.wel_flowred(flowred)   # the interval the package uses
    inside = rng.random(wn.size) < 0.75
    frac = np.where(inside, rng.uniform(0.05, 0.95, wn.size) * fr,
                    rng.uniform(fr, 1.0, wn.size))
    strt[wn] = bot[wn] + frac * (top[wn] - bot[wn])
    q = np.asarray(wel.v1, dtype=np.float64).copy()
    q[::5] = 50.0                      # injections are never reduced
    wel = dataclasses.replace(wel, v1=q, flowred=flowred)
    depth = rng.uniform(-3.0, 3.0, np.asarray(drn.nodes).size)
    depth[::5] = 0.0
    drn = dataclasses.replace(drn, depth=depth, cubic=cubic)
    return dataclasses.replace(c, packages=[wel, drn, rch], strt=strt, newton=newton,
                               newton_ur=newton)


CASES = {
    "newton": lambda: _case(True),
    "plain_linear": lambda: _case(False),
    "plain_cubic": lambda: _case(False, cubic=True),
    "newton_flowred_0": lambda: _case(True, flowred=0.0),
}


def _first_system(case):
    sim = O.GwfSimulation(case)
    perlen, nstp, steady = case.periods[0]
    sim.kper, sim.kstp = 1, 1
    sim.advance(perlen / nstp, steady)
    sim.formulate(1)
    gs = GwfSolution(case)
    gs.kper, gs.kstp = 1, 1
    gs.prepare_timestep(perlen / nstp, steady)
    amat, rhs = gs.formulate(1, want_system=True)
    return sim, gs, amat, rhs


@pytest.mark.parametrize("name", sorted(CASES))
def test_formulate_and_budget_with_options_bit_exact(name):
    case = CASES[name]()
    sim, gs, amat, rhs = _first_system(case)
    try:
        db = gs.budget()
        dev = {p.kind: gs.flow_term(p.kind) for p in case.packages}
    finally:
        gs.destroy()
    # the options change the system: without them it differs
    plain = dataclasses.replace(
        case, packages=[dataclasses.replace(p, flowred=None, depth=None, cubic=False)
                        for p in case.packages])
    sp = O.GwfSimulation(plain)
    sp.kper, sp.kstp = 1, 1
    sp.advance(case.periods[0][0] / case.periods[0][1], False)
    sp.formulate(1)
    assert not np.array_equal(sp.rhs, sim.rhs)
    bad = np.flatnonzero(amat.view(np.int64) != sim.amat.view(np.int64))
    assert bad.size == 0, (bad[:8], amat[bad[:8]], sim.amat[bad[:8]])
    bad = np.flatnonzero(rhs.view(np.int64) != sim.rhs.view(np.int64))
    assert bad.size == 0, (bad[:8], rhs[bad[:8]], sim.rhs[bad[:8]])
    ob = sim.budget()
    assert np.array_equal(db["flowja"].view(np.int64), ob["flowja"].view(np.int64))
    for k in ("WEL", "DRN"):
        assert np.array_equal(dev[k].view(np.int64), ob["simvals"][k].view(np.int64)), k
    # some withdrawals reduced (between the full rate and zero)
    wel = case.packages[0]
    q = np.asarray(wel.v1)
    red = (q < 0) & (dev["WEL"] > q) & (dev["WEL"] < 0)
    assert np.count_nonzero(red) > 0


@pytest.mark.parametrize("name", ["newton", "plain_cubic"])
def test_step_with_options_heads(name):
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
