"""The reference-side plugin: HipLinearSolverType, a LinearSolverBaseType
extension (modflow6_amd/fortran/hip_linear_solver.f90), compiled against the
reference's own LinearSolverBase / MatrixBase / SparseMatrix / VectorBase /
SeqVector / ImsLinearSettings / ConvergenceSummary modules (oracle/Makefile,
target plugin) and driven through the call sequence NumericalSolutionType
uses for such a solver (tests/fortran/hip_plugin_driver.f90): create_matrix,
create_vec_mm, the sparse builder's structure, initialize, solve,
iteration_number / is_converged.

* CPU: the library loads and runs that sequence up to the device call, which
  reports that no HIP device is present; a JA that is not the DIS structure
  is rejected by mf6hip_create.
* GPU: config 0 (3 x 100 x 100 steady confined, CG + ILU0) solved through the
  plugin reaches the oracle's FP64 ILU0-PCG heads within 10 x INNER_DVCLOSE,
  in FP64 and in mixed precision.
"""
import ctypes as C
import os

import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.ims import dis_structure
from oracle import oracle as O
from tests.helpers import oracle_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "oracle", "_ref", "libmf6hipplugin.so")

needs_plugin = pytest.mark.skipif(not os.path.exists(PLUGIN),
                                  reason="plugin library not built (needs the reference sources)")


def plugin():
    _lib.load()
    L = C.CDLL(PLUGIN)
    f = L.hip_plugin_solve
    i, d = C.c_int, C.c_double
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    f.argtypes = [i, i, i, i, ip, ip, dp, dp, dp, d, d, i, i, d, i, i,
                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
    f.restype = i
    return f


def plugin_solve(case, amat, rhs, x, precision=_lib.MF6HIP_FP64, ja=None):
    ia, ja0 = dis_structure(case.nlay, case.nrow, case.ncol)
    ja = ja0 if ja is None else ja
    ims = case.ims
    it, cv = C.c_int(), C.c_int()
    rc = plugin()(case.nlay, case.nrow, case.ncol, ja.size, ia, ja, amat, rhs, x,
                  ims.inner_dvclose, ims.inner_rclose, ims.inner_maximum, ims.ilinmeth,
                  ims.relaxation_factor, precision, 1, C.byref(it), C.byref(cv))
    return rc, it.value, cv.value


@needs_plugin
def test_plugin_runs_the_solution_call_sequence_on_cpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present: the GPU test covers the solve")
    except ImportError:
        pass
    case = configs.steady_small(nlay=3, nrow=10, ncol=12, nwel=2)
    sim, amat, rhs, x0 = oracle_system(case)
    rc, it, cv = plugin_solve(case, amat, rhs, x0.copy())
    assert rc == 2 and "no HIP device" in _lib.last_error()
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    bad = ja.copy()
    bad[[1, 2]] = bad[[2, 1]]
    rc, it, cv = plugin_solve(case, amat, rhs, x0.copy(), ja=bad)
    assert rc == 1 and "JA does not match" in _lib.last_error()


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("precision", [_lib.MF6HIP_FP64, _lib.MF6HIP_MIXED])
def test_plugin_solves_config0_like_the_oracle(precision):
    case = configs.steady_small()          # config 0: 3 x 100 x 100
    sim, amat, rhs, x0 = oracle_system(case)
    xo = x0.copy()
    icn_o, it_o, *_ = O.ImsLinear(sim.dis.ia, sim.dis.ja, sim.settings).apply(
        amat.copy(), rhs.copy(), xo)
    x = x0.copy()
    rc, it, cv = plugin_solve(case, amat.copy(), rhs.copy(), x, precision=precision)
    err = np.max(np.abs(x - xo))
    print(f"plugin precision {precision}: rc {rc}, iteration_number {it}, is_converged {cv}, "
          f"oracle iterations {it_o}, max |h - h_oracle| {err:.3e}")
    assert rc == 0, _lib.last_error()
    assert cv == 1 and icn_o == 1
    assert it > 0
    assert err < 10 * case.ims.inner_dvclose
