"""Device budget (mf6hip_gwf_budget: gwf_cq + gwf_bd) against the oracle's
restatement (oracle.GwfSimulation.budget).

The device is given the oracle's heads (and formulates at the oracle's last
formulate heads, so both hold the same npf sat): flowja must then be
bit-identical (same conductance functions, same accumulation order of the
diagonal) and the budget terms equal to 1e-12 relative (the device sums
them with fixed-shape parallel reductions, the oracle sequentially).  A
device-only run must close its own budget."""
import dataclasses

import numpy as np
import pytest

from modflow6_amd import configs
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.cases import bump_case, line_case, newton01_case

pytestmark = pytest.mark.gpu


def _transient():
    return configs.transient_newton(nlay=3, nrow=24, ncol=30, nwel=8, nper=1, nstp=2)


CASES = {
    "line": line_case,
    "newton01": newton01_case,
    "bump": bump_case,
    "steady_small": lambda: configs.steady_small(nlay=3, nrow=20, ncol=24, nwel=6),
    "transient_newton": _transient,
}


def _oracle_first_step(case):
    sim = O.GwfSimulation(case)
    perlen, nstp, steady = case.periods[0]
    sim.kper = 1
    conv, _ = sim.run_timestep(perlen / nstp, steady)
    assert conv
    return sim, perlen / nstp, steady


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_budget_matches_oracle(name):
    case = CASES[name]()
    sim, delt, steady = _oracle_first_step(case)
    ob = sim.budget()
    gs = GwfSolution(case)
    gs.kper, gs.kstp = 1, 1
    gs.prepare_timestep(delt, steady)
    import ctypes as C
    from modflow6_amd import _lib
    _lib.check(gs.L.mf6hip_gwf_set_x(gs.h, _lib.dptr(np.ascontiguousarray(sim.xtemp))), "set_x")
    gs.formulate(2)
    _lib.check(gs.L.mf6hip_gwf_set_x(gs.h, _lib.dptr(np.ascontiguousarray(sim.x))), "set_x")
    db = gs.budget()
    terms = {}
    if case.has_sto:
        terms["STO-SS"] = gs.flow_term("STO-SS")
        terms["STO-SY"] = gs.flow_term("STO-SY")
    for p in case.packages:
        terms[p.kind] = gs.flow_term(p.kind)
    gs.destroy()
    fo, fd = ob["flowja"], db["flowja"]
    bad = np.nonzero(fo.view(np.int64) != fd.view(np.int64))[0]
    assert bad.size == 0, (bad[:10], fo[bad[:10]], fd[bad[:10]])
    # the per-cell storage rates and per-bound package rates the budget file
    # holds: bit-identical as well
    want = dict(ob["simvals"])
    want["STO-SS"], want["STO-SY"] = ob["strgss"], ob["strgsy"]
    for k, v in terms.items():
        assert v.shape == want[k].shape, k
        assert np.array_equal(v.view(np.int64), want[k].view(np.int64)), k
    scale = max(max(v) for v in ob["terms"].values())
    for k, (rin, rout) in db["terms"].items():
        oin, oout = ob["terms"].get(k, (0.0, 0.0))
        assert abs(rin - oin) <= 1e-12 * scale and abs(rout - oout) <= 1e-12 * scale, k
    assert abs(db["pct_discrepancy"] - ob["pct_discrepancy"]) <= 1e-9


def test_device_run_closes_its_budget():
    """A transient Newton step solved and budgeted on the device: the terms
    balance (percent discrepancy below 1e-3 %) and the storage term is
    present."""
    case = _transient()
    gs = GwfSolution(case)
    gs.kper = 1
    perlen, nstp, steady = case.periods[0]
    gs.run_timestep(perlen / nstp, steady)
    b = gs.budget(want_flowja=False)
    gs.destroy()
    assert abs(b["pct_discrepancy"]) < 1e-3
    assert b["terms"]["STO-SS"] != (0.0, 0.0)
    assert b["terms"]["RCH"][0] > 0.0 and b["terms"]["WEL"][1] > 0.0


@pytest.mark.parametrize("name", ["steady_small", "transient_newton", "bump"])
def test_device_run_budget_within_1e6_of_oracle_run(name):
    """The device and the oracle each solve the first time step on their own
    (heads agree as the reference's own tests compare them, np.allclose);
    their budget terms then agree within 1e-6 of the largest term, the
    cell-budget bar of the comparison with the reference run."""
    case = CASES[name]()
    sim, delt, steady = _oracle_first_step(case)
    ob = sim.budget()
    gs = GwfSolution(case)
    gs.kper = 1
    gs.run_timestep(delt, steady)
    hd = gs.heads()
    db = gs.budget(want_flowja=False)
    gs.destroy()
    assert np.allclose(hd, sim.x)
    scale = max(max(v) for v in ob["terms"].values())
    for k, (rin, rout) in db["terms"].items():
        oin, oout = ob["terms"].get(k, (0.0, 0.0))
        assert abs(rin - oin) <= 1e-6 * scale and abs(rout - oout) <= 1e-6 * scale, \
            (k, rin, oin, rout, oout)


def test_device_run_writes_head_and_budget_files(tmp_path):
    """run() with a head and a budget file (OC SAVE HEAD ALL / SAVE BUDGET
    ALL): one save per time step with the step's kstp, kper, pertim and
    totim; the last save holds the final heads and the budget of them (the
    arrays themselves are pinned to the oracle bit for bit above, the
    record bytes to the reference's writers in test_binaryfile)."""
    from modflow6_amd import binaryfile as B
    case = _transient()
    gs = GwfSolution(case)
    with open(tmp_path / "m.hds", "wb") as fh, open(tmp_path / "m.cbc", "wb") as fb:
        gs.run(head_file=fh, budget_file=fb)
    hd = gs.heads()
    db = gs.budget()
    again = gs.budget_records()
    gs.destroy()
    perlen, nstp, _ = case.periods[0]
    delt = perlen / nstp
    hrec = B.read_heads(str(tmp_path / "m.hds"))
    assert len(hrec) == nstp * case.nlay
    for s in range(nstp):
        for k in range(case.nlay):
            kstp, kper, pertim, totim, text, ilay, a = hrec[s * case.nlay + k]
            assert (kstp, kper, ilay, text.strip()) == (s + 1, 1, k + 1, "HEAD")
            assert pertim == (perlen if s == nstp - 1 else delt * (s + 1))
            assert totim == delt * (s + 1)
    last = np.stack([r[6] for r in hrec[-case.nlay:]]).ravel()
    assert np.array_equal(last, hd)
    brec = B.read_budget(str(tmp_path / "m.cbc"))
    per_save = 2 + 1 + len(case.packages)   # STO-SS, STO-SY, FLOW-JA-FACE, lists
    assert len(brec) == nstp * per_save
    fj = [r for r in brec if r["text"].strip() == "FLOW-JA-FACE"]
    assert [r["kstp"] for r in fj] == list(range(1, nstp + 1))
    assert np.array_equal(fj[-1]["data"], db["flowja"])
    # the last save is the budget of the final heads, record for record
    got = open(tmp_path / "m.cbc", "rb").read()
    assert got[-len(again):] == again


def test_budget_of_odd_layout_after_dirty_device_memory():
    """Regression: the storage term is summed over every position of the
    layout, and an odd nrow * ncol layout (the 51 x 51 bump model) has
    padding slots the budget kernel never writes.  Fill and release a large
    device allocation first, so that the model's buffers are likely to land
    on memory holding nonzero values: the terms must still match the
    oracle (no storage term in this steady model)."""
    import torch
    junk = torch.full((64 * 1024 * 1024,), 12345.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    test_device_budget_matches_oracle("bump")
