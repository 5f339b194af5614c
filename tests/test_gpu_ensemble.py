"""Ensembles of uneven-size realisations solved together on the device
(mixed-precision PCG, all systems per kernel launch, HIP-graph replays)
against the oracle's FP64 ILU0-PCG solve of each realisation: heads within
10 x INNER_DVCLOSE, every realisation converged."""
import numpy as np
import pytest

from modflow6_amd.ensemble import Ensemble, realisation
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_heads(case):
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    sim.formulate(1)
    d = sim.dis
    O.lib().orc_sln_ls_fixes(sim.neq, d.ia, d.ja, sim.ibound, 1, sim.amat, sim.rhs, sim.x,
                             sim.xtemp)
    icn, it, *_ = sim.imslinear.apply(sim.amat, sim.rhs, sim.x, kstp=1)
    return sim.x, icn, it


@pytest.mark.parametrize("tune", [False, True])
def test_ensemble_matches_oracle(tune):
    cases = [realisation(s, nlay_range=(2, 5), n_range=(16, 72), nwel_range=(0, 8))
             for s in range(40, 50)]
    ens = Ensemble(cases)
    if tune:
        print(ens.autotune())
    icnvg, iters, ms = ens.solve()
    heads = [ens.heads(i) for i in range(len(cases))]
    ens.destroy()
    print("iterations", iters.tolist(), f"{ms:.1f} ms")
    for c, h, ic, it in zip(cases, heads, icnvg, iters):
        ho, icn, ito = oracle_heads(c)
        print(c.name, "device", it, "oracle", ito, "max |dh|", np.max(np.abs(h - ho)))
        assert ic == 1
        assert np.max(np.abs(h - ho)) < 10 * c.ims.inner_dvclose


def test_ensemble_solve_twice_same_result():
    cases = [realisation(s, nlay_range=(2, 3), n_range=(16, 40), nwel_range=(0, 4))
             for s in range(60, 64)]
    ens = Ensemble(cases)
    a = ens.solve()
    ha = [ens.heads(i).copy() for i in range(len(cases))]
    b = ens.solve()
    hb = [ens.heads(i) for i in range(len(cases))]
    ens.destroy()
    assert np.array_equal(a[1], b[1])
    for x, y in zip(ha, hb):
        assert np.max(np.abs(x - y)) < 10 * cases[0].ims.inner_dvclose


def test_ensemble_two_level_matches_oracle():
    """The batched solve with the aggregation coarse correction (every
    realisation its own coarse operator, set up when it is staged): heads
    within 10 x INNER_DVCLOSE of the oracle's FP64 ILU0-PCG, in fewer
    iterations than the batched one-level solve of the same systems."""
    cases = [realisation(s, nlay_range=(2, 5), n_range=(16, 72), nwel_range=(0, 8))
             for s in range(40, 50)]
    one = Ensemble(cases, keep_systems=True)
    ic1, it1, ms1 = one.solve()
    systems = one.systems
    one.destroy()
    ens = Ensemble(cases, preconditioner=2, systems=systems)
    print(ens.autotune())
    icnvg, iters, ms = ens.solve()
    heads = [ens.heads(i) for i in range(len(cases))]
    ens.destroy()
    print("one-level", it1.tolist(), f"{ms1:.1f} ms; two-level", iters.tolist(), f"{ms:.1f} ms")
    for c, h, ic in zip(cases, heads, icnvg):
        ho, icn, ito = oracle_heads(c)
        assert ic == 1
        assert np.max(np.abs(h - ho)) < 10 * c.ims.inner_dvclose
    assert iters.sum() < it1.sum()
