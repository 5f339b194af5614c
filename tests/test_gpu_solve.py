"""GPU linear solves and full simulations against the oracle.

Tolerances: heads within 10 x INNER_DVCLOSE of the oracle (FP64 natural-
order ILU0, the reference algorithm); against the oracle run on the same
column red-black permuted matrix the device follows the same Krylov
trajectory up to dot-product summation order, so iteration counts agree
within 2 % and heads within 10 x INNER_DVCLOSE; the first iterations
agree to 1e-9 relative."""
import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.ims import ConvergenceRecord
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.helpers import device_solver, oracle_permuted_factor, oracle_system

pytestmark = pytest.mark.gpu


def test_cg_solve_matches_oracle_trajectory():
    case = configs.steady_small()
    sim, amat, rhs, x0 = oracle_system(case)
    # oracle, natural order (the reference algorithm)
    xn = x0.copy()
    ims = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings)
    icn_n, it_n, _, _, _ = ims.apply(amat, rhs.copy(), xn, kstp=1)
    # oracle on the column red-black permuted matrix (the device algorithm)
    perm, (pia, pja, pa), pims, _ = oracle_permuted_factor(sim, amat)
    pb = np.empty_like(rhs); pb[perm] = rhs
    px = np.empty_like(x0); px[perm] = x0
    pims2 = O.ImsLinear(pia, pja, sim.settings)
    icn_p, it_p, _, _, recs = pims2.apply(pa, pb, px, kstp=1, history=True)
    xp = px[perm]
    # device
    sv = device_solver(case)
    xd = x0.copy()
    rec = ConvergenceRecord(case.ims.inner_maximum)
    icn_d, it_d = sv.solve(1, amat, rhs, xd, kstp=1, record=rec)
    sv.destroy()
    print(f"iterations: natural {it_n}, column-RB oracle {it_p}, device {it_d}")
    assert icn_d == 1 and icn_p == 1 and icn_n == 1
    # the exit iteration moves by a few when |r|max crosses RCLOSE at a
    # slightly different point (dot products summed in another order)
    assert abs(it_d - it_p) <= max(3, int(0.02 * it_p))
    assert np.max(np.abs(xd - xp)) < 10 * case.ims.inner_dvclose
    assert np.max(np.abs(xd - xn)) < 10 * case.ims.inner_dvclose
    # first iterations of the recorded trajectory agree closely
    for t in range(5):
        assert abs(rec.dvmax[t] - recs[t][0]) <= 1e-9 * abs(recs[t][0])
        assert abs(rec.rmax[t] - recs[t][1]) <= 1e-9 * abs(recs[t][1])


def test_bicgstab_newton_system():
    case = configs.transient_newton(seed=5, nlay=3, nrow=30, ncol=32, nwel=10,
                                    nper=1, nstp=2, perlen=10.0)
    sim, amat, rhs, x0 = oracle_system(case)
    xn = x0.copy()
    ims = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings)
    icn_n, it_n, *_ = ims.apply(amat, rhs.copy(), xn, kstp=1)
    sv = device_solver(case)
    xd = x0.copy()
    icn_d, it_d = sv.solve(1, amat, rhs, xd, kstp=1)
    sv.destroy()
    print(f"BiCGSTAB iterations: natural {it_n}, device {it_d}")
    assert icn_d == 1
    assert np.max(np.abs(xd - xn)) < 10 * case.ims.inner_dvclose


def test_simulation_steady_small_parity():
    case = configs.steady_small()
    sim = O.GwfSimulation(case)
    ho = sim.run()
    gs = GwfSolution(case)
    hd = gs.run()
    gs.destroy()
    print("oracle outer/inner:", [(r["kiter"], r["inner"]) for r in sim.log[0]])
    print("device outer/inner:", [(r.kiter, r.inner) for r in gs.log[0]])
    assert gs.log[0][-1].converged
    assert np.max(np.abs(hd - ho)) < 10 * case.ims.inner_dvclose


def test_simulation_transient_newton_parity():
    case = configs.transient_newton(seed=3, nlay=3, nrow=40, ncol=40, nwel=16,
                                    nper=2, nstp=3, perlen=20.0)
    sim = O.GwfSimulation(case)
    ho = sim.run()
    gs = GwfSolution(case)
    hd = gs.run()
    gs.destroy()
    outer_o = [len(s) for s in sim.log]
    outer_d = [len(s) for s in gs.log]
    print("outer iterations per step oracle", outer_o, "device", outer_d)
    assert all(s[-1].converged for s in gs.log)
    assert np.max(np.abs(hd - ho)) < 10 * case.ims.inner_dvclose * 10


@pytest.mark.parametrize("shape", [(3, 100, 100), (2, 33, 17), (2, 30, 30)])
@pytest.mark.parametrize("precision", [_lib.MF6HIP_MIXED, _lib.MF6HIP_MIXED_FP16])
def test_mixed_precision_steady_small(precision, shape):
    """Mixed solves on a layout where the fused kernels take two columns per
    thread (100 x 100) and on two where they take one (odd ncol; ncol / 2
    odd)."""
    nlay, nrow, ncol = shape
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol,
                                nwel=min(20, nrow * (ncol - 2) // 4))
    sim = O.GwfSimulation(case)
    ho = sim.run()
    gs = GwfSolution(case, precision=precision)
    hd = gs.run()
    gs.destroy()
    print("mixed outer/inner:", [(r.kiter, r.inner) for r in gs.log[0]],
          "max |dh|", np.max(np.abs(hd - ho)))
    assert gs.log[0][-1].converged
    assert np.max(np.abs(hd - ho)) < 10 * case.ims.inner_dvclose


@pytest.mark.parametrize("shape", [(3, 100, 100), (1, 51, 51), (5, 33, 17)])
def test_fused_cg_matches_unfused(shape):
    """The fused thread-per-column CG and the one-kernel-per-pass CG run the
    same iteration; only the time at which x / d are updated and the dot
    summation order differ."""
    nlay, nrow, ncol = shape
    case = configs.steady_small(seed=9, nlay=nlay, nrow=nrow, ncol=ncol,
                                nwel=min(20, nrow * (ncol - 2) // 4))
    sim, amat, rhs, x0 = oracle_system(case)
    res = {}
    for variant in (0, 1):
        sv = device_solver(case, kernel_variant=variant)
        xd = x0.copy()
        rec = ConvergenceRecord(case.ims.inner_maximum)
        icn, it = sv.solve(1, amat, rhs, xd, kstp=1, record=rec)
        sv.destroy()
        res[variant] = (icn, it, xd, rec)
    print("fused", res[0][1], "unfused", res[1][1])
    assert res[0][0] == 1 and res[1][0] == 1
    assert abs(res[0][1] - res[1][1]) <= max(3, int(0.02 * res[1][1]))
    assert np.max(np.abs(res[0][2] - res[1][2])) < 10 * case.ims.inner_dvclose
    for t in range(5):
        assert abs(res[0][3].dvmax[t] - res[1][3].dvmax[t]) <= 1e-9 * abs(res[1][3].dvmax[t])
        assert abs(res[0][3].rmax[t] - res[1][3].rmax[t]) <= 1e-9 * abs(res[1][3].rmax[t])


def test_fused_cg_graph_replay_matches_eager_launches():
    """Without the per-launch timing the fused FP64 CG replays captured
    eight-iteration graphs; with it, it launches every kernel.  The same
    kernels run in the same order, so heads and iteration counts agree bit
    for bit, on the first and on a repeated run of the same handle (the
    graph is reused; host synchronisations sit between the replays)."""
    case = configs.steady_small(nlay=3, nrow=40, ncol=40, nwel=8)
    eager = GwfSolution(case, timing=True)
    he = eager.run()
    inner_e = [r.inner for r in eager.log[0]]
    eager.destroy()
    gs = GwfSolution(case)
    for rep in range(2):
        if rep:
            from modflow6_amd import _lib
            _lib.check(gs.L.mf6hip_gwf_set_x(gs.h, _lib.dptr(np.ascontiguousarray(case.strt))),
                       "set_x")
            gs.log.clear()
        hg = gs.run()
        assert [r.inner for r in gs.log[0]] == inner_e
        assert np.array_equal(hg.view(np.int64), he.view(np.int64))
    gs.destroy()


@pytest.mark.parametrize("precision", [_lib.MF6HIP_MIXED, _lib.MF6HIP_MIXED_FP16])
def test_mixed_bicgstab_newton_system(precision):
    """Mixed-precision BiCGSTAB (FP32 ILU0 + SpMV BiCGSTAB inside FP64
    refinement) on a nonsymmetric Newton system: heads within 10 x
    INNER_DVCLOSE of the oracle's FP64 BiCGSTAB + ILU0 (the reference
    algorithm), FP64 true residual within INNER_RCLOSE."""
    case = configs.transient_newton(seed=5, nlay=3, nrow=30, ncol=32, nwel=10,
                                    nper=1, nstp=2, perlen=10.0)
    sim, amat, rhs, x0 = oracle_system(case)
    xn = x0.copy()
    ims = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings)
    icn_n, it_n, *_ = ims.apply(amat.copy(), rhs.copy(), xn, kstp=1)
    assert icn_n == 1
    sv = device_solver(case, precision=precision)
    xd = x0.copy()
    try:
        icn_d, it_d = sv.solve(1, amat.copy(), rhs.copy(), xd, kstp=1)
    finally:
        sv.destroy()
    d = sim.dis
    r = rhs.copy()
    ax = np.zeros_like(r)
    O.lib().orc_amux(sim.neq, xd, ax, amat, d.ja, d.ia)
    rtrue = np.max(np.abs(r - ax))
    err = np.max(np.abs(xd - xn))
    print(f"mixed BiCGSTAB precision {precision}: iterations {it_d} (oracle FP64 {it_n}), "
          f"max |h - h_oracle| {err:.3e}, max |b - A x| {rtrue:.3e}")
    assert icn_d == 1
    assert err < 10 * case.ims.inner_dvclose
    assert rtrue <= case.ims.inner_rclose


@pytest.mark.parametrize("precision", [_lib.MF6HIP_MIXED, _lib.MF6HIP_MIXED_FP16])
def test_mixed_bicgstab_transient_simulation(precision):
    """A whole transient unconfined Newton run (NEWTON UNDER_RELAXATION,
    drains, storage) with the mixed BiCGSTAB on the device-resident model:
    every step converges and the heads stay within 10 x the simulation
    tolerance of the oracle's FP64 run, as the FP64 device run does."""
    case = configs.transient_newton(seed=3, nlay=3, nrow=40, ncol=40, nwel=16,
                                    nper=2, nstp=3, perlen=20.0)
    ho = O.GwfSimulation(case).run()
    gs = GwfSolution(case, precision=precision)
    try:
        hd = gs.run()
    finally:
        gs.destroy()
    err = np.max(np.abs(hd - ho))
    print("mixed BiCGSTAB outer iterations per step", [len(s) for s in gs.log],
          "max |h - h_oracle|", err)
    assert all(s[-1].converged for s in gs.log)
    assert err < 10 * case.ims.inner_dvclose * 10
