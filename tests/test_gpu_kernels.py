"""GPU parity of the individual hot-path kernels against the oracle, bit for
bit: CSR SpMV (amux), ILU0 / MILU0 factor and apply in the column red-black
order (pcilu0 / ilu0a on the permuted matrix), and the GWF formulate
(npf_fc / sto_fc / bnd_fc / npf_fn / sto_fn + sln_ls row fixes)."""
import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.helpers import (apply_op, bits_equal, device_solver, load_factor,
                           oracle_permuted_factor, oracle_system, ulp_diff)

pytestmark = pytest.mark.gpu

SHAPES = [(3, 100, 100), (1, 51, 51), (2, 7, 9), (4, 12, 10)]


def small_case(shape, seed=7):
    nlay, nrow, ncol = shape
    return configs.steady_small(seed=seed, nlay=nlay, nrow=nrow, ncol=ncol,
                                nwel=min(20, nrow * (ncol - 2) // 4))


@pytest.mark.parametrize("shape", SHAPES)
def test_spmv_bit_exact(shape):
    case = small_case(shape)
    sim, amat, rhs, x0 = oracle_system(case)
    sv = device_solver(case)
    load_factor(sv, amat)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(sim.neq)
    y = apply_op(sv, 0, x)
    yo = np.zeros(sim.neq)
    d = sim.dis
    O.lib().orc_amux(sim.neq, x, yo, amat, d.ja, d.ia)
    assert bits_equal(y, yo), f"max ulp {ulp_diff(y, yo)}"
    sv.destroy()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("relax", [0.0, 0.97])
def test_ilu0_factor_and_apply_bit_exact(shape, relax):
    case = small_case(shape)
    case.ims.relaxation_factor = relax
    sim, amat, rhs, x0 = oracle_system(case)
    sv = device_solver(case)
    load_factor(sv, amat)
    piv = apply_op(sv, 2)
    perm, (pia, pja, pa), ims, stop = oracle_permuted_factor(sim, amat, relax)
    assert stop == -1
    piv_oracle_nat = ims.apc[:sim.neq][perm]
    assert bits_equal(piv, piv_oracle_nat), f"max ulp {ulp_diff(piv, piv_oracle_nat)}"
    rng = np.random.default_rng(2)
    r = rng.standard_normal(sim.neq)
    z = apply_op(sv, 1, r)
    pr = np.empty_like(r)
    pr[perm] = r
    pz = np.zeros_like(r)
    O.lib().orc_ilu0a(ims.nja, ims.neq, ims.apc, ims.iapc, ims.japc, pr, pz)
    zo = pz[perm]
    assert bits_equal(z, zo), f"max ulp {ulp_diff(z, zo)}"
    sv.destroy()


def test_formulate_confined_bit_exact():
    case = small_case((3, 100, 100))
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    sim.formulate(1)
    gs = GwfSolution(case)
    gs.kstp = 1
    gs.prepare_timestep(1.0, True)
    amat, rhs = gs.formulate(1, want_system=True)
    assert bits_equal(amat, sim.amat), f"amat max ulp {ulp_diff(amat, sim.amat)}"
    assert bits_equal(rhs, sim.rhs), f"rhs max ulp {ulp_diff(rhs, sim.rhs)}"
    gs.destroy()


def newton_case(shape=(3, 24, 20)):
    nlay, nrow, ncol = shape
    return configs.transient_newton(seed=11, nlay=nlay, nrow=nrow, ncol=ncol,
                                    nwel=12, nper=2, nstp=2, perlen=10.0)


def test_formulate_newton_transient_bit_exact():
    """npf_cf / npf_fn / sto_fc / sto_fn / DRN with heads taken mid-run."""
    case = newton_case()
    sim = O.GwfSimulation(case)
    perlen, nstp, steady = case.periods[0]
    sim.run_timestep(perlen / nstp, steady)          # heads move, drains flip
    sim.kstp += 1
    sim.advance(perlen / nstp, steady)
    sim.formulate(1)
    gs = GwfSolution(case)
    gs.kstp = 1
    gs.prepare_timestep(perlen / nstp, steady)
    # put the device on the oracle's state: xold from the oracle, x too
    _lib.check(_lib.load().mf6hip_gwf_set_x(gs.h, _lib.dptr(sim.xold.copy())),
               "set_x")
    gs.prepare_timestep(perlen / nstp, steady)        # xold = x (= oracle xold)
    amat, rhs = gs.formulate(1, want_system=True)
    assert bits_equal(amat, sim.amat), f"amat max ulp {ulp_diff(amat, sim.amat)}"
    assert bits_equal(rhs, sim.rhs), f"rhs max ulp {ulp_diff(rhs, sim.rhs)}"
    gs.destroy()


def test_row_fixes_bit_exact():
    case = small_case((3, 40, 30))
    sim, amat, rhs, x0 = oracle_system(case)
    gs = GwfSolution(case)
    gs.kstp = 1
    gs.prepare_timestep(1.0, True)
    gs.formulate(1)
    gs.linear_solve(1)
    amat_d, rhs_d = gs.system()
    assert bits_equal(amat_d, amat), f"amat max ulp {ulp_diff(amat_d, amat)}"
    assert bits_equal(rhs_d, rhs), f"rhs max ulp {ulp_diff(rhs_d, rhs)}"
    gs.destroy()
