"""Device path (formulate + linear solve + Newton under-relaxation, all on the
GPU) on the reference's own autotest models: closed-form heads of a 1-D CHD
line (test_par_gwf01.py:200-211), test_gwf_newton01.py (heads 8 and 7) and
the golden heads of test_gwf_newton_under_relaxation.py (results.hds.cmp,
compared with np.allclose as that test does), also with the IMS NONLINEAR
controls (backtracking, pseudo-transient continuation, under-relaxation)."""
import os

import numpy as np
import pytest

from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.cases import (BUMP_NONLINEAR, GOLD, bump_case, bump_variant, line_case,
                         newton01_case, read_ulasav)

pytestmark = pytest.mark.gpu


def test_device_linear_heads_between_two_chd():
    case = line_case()
    gs = GwfSolution(case)
    h = gs.run()
    gs.destroy()
    exact = 1.0 + 9.0 * np.arange(case.ncol) / (case.ncol - 1)
    assert np.max(np.abs(h - exact)) < 1e-9


def test_device_newton01_closed_form():
    gs = GwfSolution(newton01_case())
    h = gs.run().reshape(2, 3, 3)
    gs.destroy()
    assert np.allclose(h[0], 8.0, rtol=0, atol=1e-7)
    assert np.allclose(h[1], 7.0, rtol=0, atol=1e-12)


def test_device_bump_newton_under_relaxation_golden():
    case = bump_case()
    gs = GwfSolution(case)
    h = gs.run().reshape(51, 51)
    recs = gs.log[0]
    gs.destroy()
    sim = O.GwfSimulation(case)
    ho = sim.run().reshape(51, 51)
    print("device outer iterations", len(recs), "oracle", len(sim.log[0]),
          "NUR engaged", sum(r.nur for r in recs))
    base = read_ulasav(os.path.join(GOLD, "ex-gwf-bump", "results.hds.cmp"))
    assert recs[-1].converged
    assert any(r.nur for r in recs)
    assert np.allclose(base, h)
    assert np.max(np.abs(h - ho)) < 1e-6


@pytest.mark.parametrize("name", sorted(BUMP_NONLINEAR))
def test_device_nonlinear_controls(name):
    """Backtracking, PTC and SIMPLE / COOLEY / DBD under-relaxation on the
    device: the golden heads, the oracle's heads to 1e-6 m, and the oracle's
    outer-iteration path (largest head change, PTC applied) over the first
    iterations.  Later outer iterations may part ways: DBD and the
    backtracking test compare quantities whose sign or threshold flips on
    round-off once the heads are nearly converged, so only the end result and
    the early path are compared."""
    case = bump_variant(name)
    gs = GwfSolution(case)
    h = gs.run().reshape(51, 51)
    recs = gs.log[0]
    gs.destroy()
    sim = O.GwfSimulation(case)
    ho = sim.run().reshape(51, 51)
    orecs = sim.log[0]
    print(name, "device outer", len(recs), "oracle", len(orecs),
          "bt", sum(r.nbt for r in recs), sum(r["nbt"] for r in orecs),
          "ptc", sum(r.iptc for r in recs), sum(r["iptc"] for r in orecs))
    base = read_ulasav(os.path.join(GOLD, "ex-gwf-bump", "results.hds.cmp"))
    assert recs[-1].converged
    assert np.allclose(base, h)
    assert np.max(np.abs(h - ho)) < 1e-6
    # inner solves stop at INNER_DVCLOSE 1e-9 with dot products summed in
    # another order, so each outer iterate carries differences of that order
    # and more; the early path agrees to 1e-4 relative
    for rd, ro in list(zip(recs, orecs))[:4]:
        assert abs(rd.hncg - ro["hncg"]) <= 1e-7 + 1e-4 * abs(ro["hncg"])
        assert rd.iptc == ro["iptc"] and rd.nbt == ro["nbt"]
    if name in ("ptc", "complex"):
        assert sum(r.iptc for r in recs) > 0


def test_device_backtrack_and_residual_norm():
    """mf6hip_gwf_backtrack moves the heads to xtemp + breduc (x - xtemp)
    (apply_backtracking, bit for bit) and mf6hip_gwf_residual_norms equals
    the oracle's sln_l2norm of the same system."""
    case = bump_variant("complex")
    gs = GwfSolution(case)
    gs.kper, gs.kstp = 1, 1
    gs.prepare_timestep(1.0, True)
    x0 = gs.heads()
    gs.solve(1)
    x1 = gs.heads()
    amat, rhs = gs.formulate(2, want_system=True, inewton=0)
    l2, rv = gs.residual_norms()
    import ctypes as C
    from modflow6_amd import _lib
    flag = C.c_int(0)
    _lib.check(gs.L.mf6hip_gwf_backtrack(gs.h, 0.5, 1e-8, C.byref(flag)), "backtrack")
    x2 = gs.heads()
    gs.destroy()
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    active = sim.ibound > 0
    assert flag.value == 1
    exp = x1.copy()
    exp[active] = x0[active] + 0.5 * (x1[active] - x0[active])
    assert np.array_equal(x2.view(np.int64), exp.view(np.int64))
    d = sim.dis
    r = np.zeros(sim.neq)
    lo = O.lib().orc_sln_l2norm(sim.neq, d.ia, d.ja, amat, x1, rhs, sim.ibound, r)
    assert abs(l2 - lo) <= 1e-12 * lo
    vol = d.area * (d.top - d.bot)
    assert abs(rv - np.max(np.abs(r[active]) / vol[active])) <= 1e-12 * rv


def test_device_head_exchange_with_device_buffers():
    """mf6hip_gwf_get_x_device / set_x_device move heads between the model and
    a caller's GPU buffer (DIS node order) without a host copy: bit-identical
    to the host exchange, and a solve started from heads set this way ends
    where the host-set start ends."""
    import torch
    from modflow6_amd import _lib
    case = bump_case()
    gs = GwfSolution(case)
    gs.kper, gs.kstp = 1, 1
    gs.prepare_timestep(1.0, True)
    gs.solve(1)
    xh = gs.heads()
    buf = torch.full((case.nodes,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(gs.L.mf6hip_gwf_get_x_device(gs.h, buf.data_ptr()), "get_x_device")
    xd = buf.cpu().numpy()
    assert np.array_equal(xd.view(np.int64), xh.view(np.int64))
    # perturb on the device, set back, read through the host path
    buf.add_(0.25)
    torch.cuda.synchronize()
    _lib.check(gs.L.mf6hip_gwf_set_x_device(gs.h, buf.data_ptr()), "set_x_device")
    xs = gs.heads()
    assert np.array_equal(xs.view(np.int64), (xh + 0.25).view(np.int64))
    gs.destroy()
