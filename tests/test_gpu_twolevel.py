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
    # odd ncol: the single-column kernels and the odd-ncol place of a column
    # in the i-edge coupling rows of the deflated form
    "steady_odd": lambda: configs.steady_small(nlay=2, nrow=13, ncol=17, nwel=3),
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


# ---------------------------------------------------------------------------
# The operators themselves against a sparse numpy build (no dense Q)
# ---------------------------------------------------------------------------

def dyadic_dis_matrix(shape, seed, chd_columns=True):
    """A symmetric, negative definite DIS-structured matrix whose entries
    (and every partial sum of a row) are exact in FP32: off-diagonals k/8,
    k in 1..24 (5 % zero), diagonals minus the row's sum minus j/8.  With
    chd_columns the first and last grid columns are decoupled identity-like
    rows (CHD cells after the sln_ls fixes)."""
    from modflow6_amd.ims import dis_structure
    nlay, nrow, ncol = shape
    ia, ja = dis_structure(nlay, nrow, ncol)
    ia0 = ia.astype(np.int64) - 1
    ja0 = ja.astype(np.int64) - 1
    n = nlay * nrow * ncol
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), np.diff(ia0))
    up = ja0 > rows
    vals = rng.integers(1, 25, up.sum()) / 8.0
    vals[rng.random(up.sum()) < 0.05] = 0.0
    a = np.zeros(ia0[-1])
    a[up] = vals
    if chd_columns:
        jcol = np.arange(n) % ncol
        fixed = (jcol == 0) | (jcol == ncol - 1)
        a[up & (fixed[rows] | fixed[ja0])] = 0.0
    import scipy.sparse as sp
    U = sp.csr_matrix((a, ja0, ia0), shape=(n, n))
    S = (U + U.T).tocsr()
    S.sort_indices()
    lo = ja0 < rows
    a[lo] = np.asarray(S[rows[lo], ja0[lo]]).ravel()
    diag = ia0[:-1]
    a[diag] = -(np.bincount(rows, weights=a, minlength=n) + rng.integers(0, 9, n) / 8.0) - 0.125
    return ia, ja, a


def sparse_coarse(shape, ia, ja, a, target):
    """A, P and P^T A P as scipy sparse matrices (natural node order);
    aggregates of ha x ha grid columns, P restricted to coupled rows, empty
    aggregates given a -1 diagonal."""
    import scipy.sparse as sp
    nlay, nrow, ncol = shape
    n = nlay * nrow * ncol
    A = sp.csr_matrix((a, ja - 1, ia - 1), shape=(n, n))
    off = A - sp.diags(A.diagonal())
    coupled = np.asarray((off != 0).sum(axis=1)).ravel() > 0
    ha, nai, naj = coarse_shape(nrow, ncol, target)
    k, i, j = np.unravel_index(np.arange(n), (nlay, nrow, ncol))
    agg = (i // ha) * naj + j // ha
    nc = nai * naj
    r = np.flatnonzero(coupled)
    P = sp.csr_matrix((np.ones(r.size), (r, agg[r])), shape=(n, nc))
    Ac = (P.T @ A @ P).tocsc()
    empty = np.asarray(P.sum(axis=0)).ravel() == 0
    Ac = Ac + sp.diags(np.where(empty, -1.0, 0.0))
    return A, P, Ac.tocsc(), ha, naj


DEFL_CASES = {
    "3x20x24": ((3, 20, 24), 40),
    "odd_2x13x17": ((2, 13, 17), 12),
    "one_column_2x6x8": ((2, 6, 8), 64),
    "4x30x40": ((4, 30, 40), 100),
}


@pytest.mark.parametrize("name", sorted(DEFL_CASES))
def test_deflated_operator_matches_sparse_build(name):
    """op 6 (the deflated preconditioner of the fused CG: yc = Ac^-1 P^T r,
    z = M^-1 (r - A P yc) + P yc) against yc from a sparse solve of
    P^T A P, A P yc from the sparse product, and M^-1 from the device's ILU0
    apply (op 1, bit-exact against the oracle's ilu0a).  The matrix entries
    are exact in FP32, so the FP32 row sums and edge couplings the sweeps
    read are exact and the only differences are the FP64 roundings of the
    two coarse solves: tolerance 1e-9 of the largest |z|.  A wrong place
    in the edge-coupling rows moves z by O(1)."""
    import scipy.sparse.linalg as spla
    shape, target = DEFL_CASES[name]
    ia, ja, a = dyadic_dis_matrix(shape, seed=sum(shape))
    assert np.array_equal(a.astype(np.float32).astype(np.float64), a)
    A, P, Ac, ha, naj = sparse_coarse(shape, ia, ja, a, target)
    if name.startswith("one_column"):
        assert ha == 1
    nlay, nrow, ncol = shape
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol, nwel=0)
    sv = device_solver(case, preconditioner=3, coarse_size=target)
    try:
        load_factor(sv, a)
        rng = np.random.default_rng(4)
        for _ in range(3):
            r = rng.standard_normal(case.nodes)
            yc = spla.spsolve(Ac, P.T @ r)
            w = r - A @ (P @ yc)
            want = apply_op(sv, 1, w) + P @ yc
            got = apply_op(sv, 6, r)
            err = np.max(np.abs(got - want))
            print(name, "ha", ha, "max |z - z_sparse|", err, "scale", np.max(np.abs(want)))
            assert err <= 1e-9 * np.max(np.abs(want))
    finally:
        sv.destroy()


def test_coarse_operator_63_wide_aggregate_grid():
    """The additive coarse part (op 3 - op 1) on a 63 x 63 aggregate grid
    (nc = 3969, the widest the banded coarse Cholesky takes, as at the 10M
    cell headline): P (P^T A P)^-1 P^T x from a sparse solve, without a dense
    Q.  Tolerance 1e-9 of the largest value."""
    import scipy.sparse.linalg as spla
    shape, target = (2, 252, 252), 3969
    ia, ja, a = dyadic_dis_matrix(shape, seed=9)
    A, P, Ac, ha, naj = sparse_coarse(shape, ia, ja, a, target)
    assert (ha, naj, P.shape[1]) == (4, 63, 3969)
    nlay, nrow, ncol = shape
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol, nwel=0)
    sv = device_solver(case, preconditioner=2, coarse_size=target)
    try:
        load_factor(sv, a)
        rng = np.random.default_rng(6)
        for _ in range(2):
            x = rng.standard_normal(case.nodes)
            got = apply_op(sv, 3, x) - apply_op(sv, 1, x)
            want = P @ spla.spsolve(Ac, P.T @ x)
            err = np.max(np.abs(got - want))
            print("63-wide: max |coarse - sparse|", err, "scale", np.max(np.abs(want)))
            assert err <= 1e-9 * np.max(np.abs(want))
    finally:
        sv.destroy()


GATHER_CASES = {
    "3x20x24": ((3, 20, 24), 40),
    "odd_2x13x17": ((2, 13, 17), 12),
    "one_column_2x6x8": ((2, 6, 8), 64),
    # 1000 columns, ha = 16: rows of column sums not aligned to cache lines,
    # the last aggregate of each row cut by the edge, four column chunks
    "2x40x1000": ((2, 40, 1000), 189),
}


@pytest.mark.parametrize("name", sorted(GATHER_CASES))
def test_deflated_restriction_gather(name):
    """op 9: the restriction the deflated iteration forms from column sums
    (P^T (r - alpha q) from the sweeps' natural-order sums of r and the
    SpMV's colour-order sums of q; here r = q = x and alpha = 0.75) against
    the direct restriction over the coupled cells (k_tl_restrict) and
    against P^T x from the sparse P (coupled rows only).  Tolerance 1e-12
    of the largest |P^T x| (the three sum in different orders)."""
    shape, target = GATHER_CASES[name]
    ia, ja, a = dyadic_dis_matrix(shape, seed=sum(shape) + 1)
    A, P, Ac, ha, naj = sparse_coarse(shape, ia, ja, a, target)
    nc = P.shape[1]
    nlay, nrow, ncol = shape
    case = configs.steady_small(nlay=nlay, nrow=nrow, ncol=ncol, nwel=0)
    sv = device_solver(case, preconditioner=3, coarse_size=target)
    try:
        load_factor(sv, a)
        rng = np.random.default_rng(5)
        for _ in range(2):
            x = rng.standard_normal(case.nodes)
            y = apply_op(sv, 9, x)
            want = P.T @ x
            scale = np.max(np.abs(want))
            print(name, "ha", ha, "nc", nc, "gather err", np.max(np.abs(y[:nc] - 0.25 * want)),
                  "restrict err", np.max(np.abs(y[nc:2 * nc] - want)), "scale", scale)
            assert np.max(np.abs(y[nc:2 * nc] - want)) <= 1e-12 * scale
            assert np.max(np.abs(y[:nc] - 0.25 * want)) <= 1e-12 * scale
            assert not np.any(y[2 * nc:])
    finally:
        sv.destroy()
