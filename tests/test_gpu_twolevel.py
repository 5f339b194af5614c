"""Two-level preconditioner (preconditioner = 2): ILU0 plus the additive
coarse correction P (P^T A P)^-1 P^T on column aggregates.

* The coarse part of the device operator equals P (P^T A P)^-1 P^T x built
  here with numpy from the CSR matrix: the aggregates of ha x ha grid
  columns (all layers), P restricted to the rows with a nonzero
  off-diagonal (CHD rows after the sln_ls fixes are identity rows and are
  left out), a dense solve of the coarse system.  Measured: the device's
  two-level apply minus its one-level apply (op 3 - op 1), so the ILU0 part
  cancels exactly.  Tolerance 1e-9 of the largest value (the device
  inverts the coarse matrix with Cholesky, numpy solves it with LU).
* Solves in FP64, MIXED and MIXED_FP16 give heads within 10 x inner_dvclose
  of the oracle's FP64 ILU0-PCG (the reference's own algorithm) and take
  fewer iterations than that solve.
"""
import numpy as np
import pytest

from modflow6_amd import _lib, configs
from oracle import oracle as O
from tests.helpers import apply_op, device_solver, load_factor, oracle_system

pytestmark = pytest.mark.gpu


def coarse_shape(nrow, ncol, target):
    """Smallest aggregate width ha with ceil(nrow/ha) ceil(ncol/ha) <= target,
    widened (at most to twice that) until ceil(ncol/ha) <= 63."""
    target = target or 4096
    ha = max(1, int(np.ceil(np.sqrt(nrow * ncol / target))))
    while -(-nrow // ha) * -(-ncol // ha) > target:
        ha += 1
    ha0 = ha
    while -(-ncol // ha) > 63 and ha < 2 * ha0:
        ha += 1
    return ha, -(-nrow // ha), -(-ncol // ha)


def coarse_operator(case, amat, ia, ja, target):
    """Dense P (P^T A P)^-1 P^T (natural node order)."""
    nlay, nrow, ncol = case.nlay, case.nrow, case.ncol
    n = nlay * nrow * ncol
    ha, nai, naj = coarse_shape(nrow, ncol, target)
    coupled = np.zeros(n, dtype=bool)
    for r in range(n):
        off = amat[ia[r] + 1:ia[r + 1]]      # 0-based ia, diagonal first
        coupled[r] = np.any(off != 0.0)
    k, i, j = np.unravel_index(np.arange(n), (nlay, nrow, ncol))
    agg = (i // ha) * naj + j // ha
    nc = nai * naj
    P = np.zeros((n, nc))
    P[np.arange(n)[coupled], agg[coupled]] = 1.0
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ja[ia[r]:ia[r + 1]]] = amat[ia[r]:ia[r + 1]]
    Ac = P.T @ A @ P
    empty = ~np.any(P != 0.0, axis=0)
    Ac[empty, empty] = -1.0
    return P @ np.linalg.solve(Ac, P.T), ha


CASES = {
    # CHD on two grid columns, wells, recharge
    "steady_small": (lambda: configs.steady_small(nlay=3, nrow=20, ncol=24, nwel=6), 40),
    # odd ncol (single-column direction update), aggregates cut by the edge
    "steady_odd": (lambda: configs.steady_small(nlay=2, nrow=13, ncol=17, nwel=3), 12),
    # an aggregate grid 150 wide: the dense (rocSOLVER) coarse factorisation
    "steady_wide": (lambda: configs.steady_small(nlay=2, nrow=4, ncol=300, nwel=3), 1200),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_coarse_operator_matches_dense_solve(name):
    mk, target = CASES[name]
    case = mk()
    sim, amat, rhs, x0 = oracle_system(case)
    d = sim.dis
    Q, ha = coarse_operator(case, amat, d.ia, d.ja, target)
    if name == "steady_wide":
        assert -(-case.ncol // ha) > 63     # wider than the banded setup takes
    sv = device_solver(case, preconditioner=2, coarse_size=target)
    try:
        load_factor(sv, amat)
        rng = np.random.default_rng(3)
        for _ in range(3):
            x = rng.standard_normal(case.nodes)
            got = apply_op(sv, 3, x) - apply_op(sv, 1, x)
            want = Q @ x
            err = np.max(np.abs(got - want))
            print(name, "ha", ha, "max |coarse - dense|", err, "scale", np.max(np.abs(want)))
            assert err <= 1e-9 * np.max(np.abs(want))
    finally:
        sv.destroy()


def _oracle_heads(sim, amat, rhs, x0):
    x = x0.copy()
    icnvg, its, *_ = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings).apply(
        amat.copy(), rhs.copy(), x)
    assert icnvg == 1
    return x, its


SOLVE_CASES = {
    "steady_small": lambda: configs.steady_small(nlay=3, nrow=20, ncol=24, nwel=6),
    "steady_60": lambda: configs.steady_10m(nlay=4, nrow=60, ncol=64, nwel=12),
}


@pytest.mark.parametrize("precision", [_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED,
                                       _lib.MF6HIP_MIXED_FP16])
@pytest.mark.parametrize("name", sorted(SOLVE_CASES))
def test_twolevel_solve_heads(name, precision):
    case = SOLVE_CASES[name]()
    sim, amat, rhs, x0 = oracle_system(case)
    xo, its_o = _oracle_heads(sim, amat, rhs, x0)
    sv = device_solver(case, precision=precision, preconditioner=2, coarse_size=64)
    x = x0.copy()
    try:
        icnvg, its = sv.solve(1, amat.copy(), rhs.copy(), x)
    finally:
        sv.destroy()
    err = np.max(np.abs(x - xo))
    print(name, "precision", precision, "oracle ILU0-PCG iterations", its_o,
          "two-level iterations", its, "max |h - h_oracle|", err)
    assert icnvg == 1, its
    assert err <= 10 * case.ims.inner_dvclose
    assert its < its_o


def test_twolevel_rejects_unsymmetric_paths():
    case = configs.steady_small(nlay=2, nrow=6, ncol=8, nwel=2)
    with pytest.raises(RuntimeError):
        device_solver(case, preconditioner=2, kernel_variant=1)


@pytest.mark.parametrize("precision", [_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED_FP16])
def test_twolevel_outer_loop_matches_oracle(precision):
    """Whole steady simulation on the device-resident model (formulate, row
    fixes, two-level solve, outer updates each outer iteration) against the
    oracle's run of the reference algorithm."""
    from modflow6_amd.solution import GwfSolution
    case = configs.steady_10m(nlay=3, nrow=48, ncol=50, nwel=6)
    ho = O.GwfSimulation(case).run()
    gs = GwfSolution(case, precision=precision, preconditioner=2, coarse_size=100)
    try:
        hd = gs.run()
        log = gs.log[0]
    finally:
        gs.destroy()
    err = np.max(np.abs(hd - ho))
    print("outer/inner", [(r.kiter, r.inner) for r in log], "max |h - h_oracle|", err)
    assert log[-1].converged
    assert err < 10 * case.ims.inner_dvclose


@pytest.mark.parametrize("precision", [_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED,
                                       _lib.MF6HIP_MIXED_FP16])
@pytest.mark.parametrize("name", sorted(SOLVE_CASES))
def test_deflated_solve_heads(name, precision):
    """Deflated form (preconditioner = 3: z = M^-1 (r - A P yc) + P yc, the
    sweeps of the fused CG fed r - A P yc): heads within 10 x inner_dvclose
    of the oracle's FP64 ILU0-PCG, in fewer iterations than it, and in no
    more than the additive form's (same coarse space) plus a few."""
    case = SOLVE_CASES[name]()
    sim, amat, rhs, x0 = oracle_system(case)
    xo, its_o = _oracle_heads(sim, amat, rhs, x0)
    its_by = {}
    for pc in (2, 3):
        sv = device_solver(case, precision=precision, preconditioner=pc, coarse_size=64)
        x = x0.copy()
        try:
            icnvg, its = sv.solve(1, amat.copy(), rhs.copy(), x)
        finally:
            sv.destroy()
        its_by[pc] = its
        err = np.max(np.abs(x - xo))
        print(name, "precision", precision, "preconditioner", pc, "iterations", its,
              "oracle", its_o, "max |h - h_oracle|", err)
        assert icnvg == 1, its
        assert err <= 10 * case.ims.inner_dvclose
        assert its < its_o
    assert its_by[3] <= its_by[2] + 3


@pytest.mark.parametrize("precision", [_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED_FP16])
def test_deflated_outer_loop_matches_oracle(precision):
    from modflow6_amd.solution import GwfSolution
    case = configs.steady_10m(nlay=3, nrow=48, ncol=50, nwel=6)
    ho = O.GwfSimulation(case).run()
    gs = GwfSolution(case, precision=precision, preconditioner=3, coarse_size=100)
    try:
        hd = gs.run()
        log = gs.log[0]
    finally:
        gs.destroy()
    err = np.max(np.abs(hd - ho))
    print("outer/inner", [(r.kiter, r.inner) for r in log], "max |h - h_oracle|", err)
    assert log[-1].converged
    assert err < 10 * case.ims.inner_dvclose


@pytest.mark.parametrize("pc", [2, 3])
def test_one_column_aggregates(pc):
    """coarse_size at least nrow x ncol: every aggregate is one grid column
    (ha = 1), so each column sits on both edges in each direction; both
    forms still reach the oracle's heads."""
    case = configs.steady_small(nlay=2, nrow=6, ncol=8, nwel=2)
    assert coarse_shape(case.nrow, case.ncol, 64)[0] == 1
    sim, amat, rhs, x0 = oracle_system(case)
    xo, its_o = _oracle_heads(sim, amat, rhs, x0)
    for precision in (_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED_FP16):
        sv = device_solver(case, precision=precision, preconditioner=pc, coarse_size=64)
        x = x0.copy()
        try:
            icnvg, its = sv.solve(1, amat.copy(), rhs.copy(), x)
        finally:
            sv.destroy()
        err = np.max(np.abs(x - xo))
        print("preconditioner", pc, "precision", precision, "iterations", its, "oracle", its_o,
              "max |h - h_oracle|", err)
        assert icnvg == 1, its
        assert err <= 10 * case.ims.inner_dvclose
