"""GPU parity of the individual hot-path kernels against the oracle, bit for
bit: CSR SpMV (amux), ILU0 / MILU0 factor and apply in the column red-black
order (pcilu0 / ilu0a on the permuted matrix), and the GWF formulate
(npf_fc / sto_fc / bnd_fc / npf_fn / sto_fn + sln_ls row fixes)."""
import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.helpers import (apply_op, bits_equal, device_solver, dis_matrix, f32_amux,
                           half_scale_exp, load_factor, oracle_permuted_factor, oracle_system,
                           rb_ilu_apply, ulp_diff)

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


# Fused-CG SpMV (ops 4 and 5): shapes that take each kernel form.  FP32
# groups 4 columns per thread when ncol % 8 == 0, 2 when ncol % 4 == 0;
# FP64 groups 2 when ncol % 4 == 0; other shapes take the single-column
# kernel.  Layer counts cover each unrolled bound (4, 8, 16, 32).
FUSED_SHAPES = [(3, 100, 100), (2, 16, 24), (10, 40, 48), (1, 51, 51), (2, 7, 9),
                (4, 12, 10), (5, 30, 40), (17, 8, 16), (3, 9, 8), (6, 3, 8)]


@pytest.mark.parametrize("shape", FUSED_SHAPES)
def test_fused_spmv_bit_exact(shape):
    nlay, nrow, ncol = shape
    ia, ja, a = dis_matrix(shape, seed=nlay * 1000 + ncol)
    n = ia.size - 1
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol, nwel=0)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    yo = np.zeros(n)
    O.lib().orc_amux(n, x, yo, a, ja - 1, ia - 1)
    sv = device_solver(case)
    try:
        load_factor(sv, a)
        y = apply_op(sv, 4, x)
        assert bits_equal(y, yo), f"FP64 max ulp {ulp_diff(y, yo)}"
    finally:
        sv.destroy()
    sv = device_solver(case, precision=_lib.MF6HIP_MIXED)
    try:
        load_factor(sv, a)
        y32 = apply_op(sv, 5, x)
        want = f32_amux(ia, ja, a, x)
        assert bits_equal(y32, want), f"FP32 max ulp {ulp_diff(y32, want)}"
    finally:
        sv.destroy()



@pytest.mark.parametrize("shape", FUSED_SHAPES)
def test_fused_ilu_sweeps_bit_exact(shape):
    """The fused CG's ILU0 sweeps (grouped kernels where ncol allows, else
    the single-column ones): in FP64 (op 7) equal to the oracle's ilu0a on
    the permuted matrix (through op 1, itself bit-exact against it); in FP32
    (op 8, MIXED) and with the scaled FP16 planes (op 8, MIXED_FP16) equal
    to the numpy colour-sweep restatement in that precision, which the CPU
    suite pins to ilu0a in FP64."""
    nlay, nrow, ncol = shape
    ia, ja, a = dis_matrix(shape, seed=nlay * 100 + nrow + ncol)
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol, nwel=0)
    x = np.random.default_rng(8).standard_normal(case.nodes)
    sv = device_solver(case)
    try:
        load_factor(sv, a)
        z1 = apply_op(sv, 1, x)
        z7 = apply_op(sv, 7, x)
        piv = apply_op(sv, 2)
        assert bits_equal(z7, z1), f"FP64 max ulp {ulp_diff(z7, z1)}"
    finally:
        sv.destroy()
    k = half_scale_exp(shape, ia, ja, a)
    for precision, hk in ((_lib.MF6HIP_MIXED, None), (_lib.MF6HIP_MIXED_FP16, k)):
        sv = device_solver(case, precision=precision)
        try:
            load_factor(sv, a)
            z8 = apply_op(sv, 8, x)
        finally:
            sv.destroy()
        want = rb_ilu_apply(shape, ia, ja, a, piv, x, dtype=np.float32, half_scale_exp=hk)
        assert bits_equal(z8, want), f"precision {precision}: max ulp {ulp_diff(z8, want)}"
