"""NPF options on the device against the oracle's NPF (pinned to the
reference's npf_fc / npf_fn / npf_cq with VARIABLECV, DEWATERED and the
alternative cell averagings in tests/test_oracle_gwf_vs_reference.py and
tests/test_oracle_vs_reference.py; K22 in tests/test_oracle_npf_k22.py):
VARIABLECV with and without DEWATERED, under NEWTON and without,
ALTERNATIVE_CELL_AVERAGING LOGARITHMIC / AMT-LMK / AMT-HMK, K22 on the
row-direction faces, and ANGLE1 (with ANGLE2 / ANGLE3) rotating the
conductivity tensor (hyeff computed on the host at create, as the oracle
does, so the faces match bit for bit), and the STO options
STORAGECOEFFICIENT, SS_CONFINED_ONLY and the original specific storage
formulation (sto_fc / sto_fn / sto_cq, pinned with the same flags), and
THICKSTRT: the formulated system and the budget's face flows bit for
bit (to 1e-14 relative where a logarithmic mean calls log()), and the heads
of solved steps within 10 x dvclose."""
import dataclasses

import numpy as np
import pytest

from modflow6_amd import configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _convertible(newton, ivarcv, idewatcv, seed=6):
    """All cells convertible, heads spread through the upper layer's
    thickness (some cells dewatered) over a deeper water table."""
    c = configs.transient_newton(seed=seed, nlay=3, nrow=24, ncol=30, nwel=8, nper=1, nstp=2)
    rng = np.random.default_rng(seed)
    bot = c.botm.reshape(-1)
    top = np.concatenate([c.top2d.reshape(-1), c.botm[:-1].reshape(-1)])
    strt = bot + rng.uniform(0.15, 0.98, bot.size) * (top - bot)
    return dataclasses.replace(c, strt=strt, newton=newton, newton_ur=newton, ivarcv=ivarcv,
                               idewatcv=idewatcv)


def _averaging(icellavg):
    c = configs.steady_small(nlay=3, nrow=20, ncol=24, nwel=6)
    return dataclasses.replace(c, icellavg=icellavg)


# the averagings with a logarithmic mean (condmean, GwfConductanceUtils.f90:
# 226-305)
LOG_MEANS = {"logarithmic", "amt_lmk"}

def _k22(base, seed=8):
    c = base()
    rng = np.random.default_rng(seed)
    return dataclasses.replace(c, k22=np.asarray(c.k11) * rng.lognormal(-1.0, 0.7, c.nodes))


def _angles(base, vertical, seed=9):
    c = _k22(base, seed)
    rng = np.random.default_rng(seed)
    a1 = rng.uniform(-90.0, 90.0, c.nodes)
    if not vertical:
        return dataclasses.replace(c, angle1=a1)
    return dataclasses.replace(c, angle1=a1, angle2=rng.uniform(-30.0, 30.0, c.nodes),
                               angle3=rng.uniform(-30.0, 30.0, c.nodes))


def _thickstrt(newton, seed=10):
    """A third of the cells entered with ICELLTYPE -1 under THICKSTRT:
    confined, with the saturated thickness of their starting head."""
    c = _convertible(newton, 0, 0, seed)
    rng = np.random.default_rng(seed)
    ic = np.asarray(c.icelltype).copy()
    ic[rng.random(c.nodes) < 0.33] = -1
    return dataclasses.replace(c, icelltype=ic, thickstrt=True)


CASES = {
    "varcv_newton": lambda: _convertible(True, 1, 0),
    "varcv_dewatered_newton": lambda: _convertible(True, 1, 1),
    "varcv_plain": lambda: _convertible(False, 1, 0),
    "varcv_dewatered_plain": lambda: _convertible(False, 1, 1),
    "logarithmic": lambda: _averaging(1),
    "amt_lmk": lambda: _averaging(2),
    "amt_hmk": lambda: _averaging(3),
    "k22_confined": lambda: _k22(lambda: _averaging(0)),
    "k22_newton": lambda: _k22(lambda: _convertible(True, 0, 0)),
    "angle1_confined": lambda: _angles(lambda: _averaging(0), False),
    "angles_newton": lambda: _angles(lambda: _convertible(True, 0, 0), True),
    "sto_storagecoefficient": lambda: dataclasses.replace(_convertible(True, 0, 0),
                                                          istor_coef=1),
    "sto_confined_only": lambda: dataclasses.replace(_convertible(True, 0, 0), iconf_ss=1),
    "sto_original": lambda: dataclasses.replace(_convertible(True, 0, 0), iorig_ss=1),
    "sto_confined_only_plain": lambda: dataclasses.replace(_convertible(False, 0, 0),
                                                           iconf_ss=1),
    "thickstrt_newton": lambda: _thickstrt(True),
    "thickstrt_plain": lambda: _thickstrt(False),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_formulate_and_budget_bit_exact(name):
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
    base = O.GwfSimulation(dataclasses.replace(case, ivarcv=0, idewatcv=0, icellavg=0,
                                               k22=None, angle1=None, angle2=None,
                                               angle3=None, istor_coef=0, iconf_ss=0,
                                               iorig_ss=0, thickstrt=False))
    base.kper, base.kstp = 1, 1
    base.advance(perlen / nstp, steady)
    base.formulate(1)
    assert not (np.array_equal(base.amat, sim.amat) and np.array_equal(base.rhs, sim.rhs))
    ob = sim.budget()
    if name in LOG_MEANS:
        # logarithmic means call log(): the device's and the host's libm
        # may differ in the last bit, so the conductances agree to a few
        # units in the last place rather than bit for bit
        np.testing.assert_allclose(amat, sim.amat, rtol=1e-14, atol=0)
        np.testing.assert_allclose(rhs, sim.rhs, rtol=1e-14, atol=0)
        fmax = np.max(np.abs(ob["flowja"]))
        np.testing.assert_allclose(db["flowja"], ob["flowja"], rtol=0, atol=1e-13 * fmax)
        return
    bad = np.flatnonzero(amat.view(np.int64) != sim.amat.view(np.int64))
    assert bad.size == 0, (bad[:8], amat[bad[:8]], sim.amat[bad[:8]])
    assert np.array_equal(rhs.view(np.int64), sim.rhs.view(np.int64))
    assert np.array_equal(db["flowja"].view(np.int64), ob["flowja"].view(np.int64))


@pytest.mark.parametrize("name", ["varcv_dewatered_newton", "logarithmic", "k22_newton",
                                  "angles_newton", "sto_storagecoefficient",
                                  "sto_confined_only", "thickstrt_newton"])
def test_step_heads(name):
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
