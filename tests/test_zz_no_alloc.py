"""mf6hip_csr_residual_norms_noalloc: the residual norms of
mf6hip_gwf_residual_norms (r = A x - b on the active rows: ||r||_2 and
||b||_2, sln_calc_residual + sln_l2norm) for a system the caller holds on the
device in the reference's CSR layout, with the library's partial sums in a
fixed device array instead of an allocation.  Checked through the C-ABI
against the existing entry point on the same system (the device-resident
GWF model's formulated system, copied out and handed over as device
arrays): ||r||_2 within 1e-12 relative (the two sum in different orders),
the signed largest r and its row bit for bit against the oracle's amux.
Two calls in a row, the second on a smaller system, so the second must not
see the first call's partials."""
import ctypes as C

import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.ims import dis_structure
from modflow6_amd.solution import GwfSolution
from oracle import oracle as O
from tests.helpers import HipBuffers

pytestmark = pytest.mark.gpu


def _device_system(case):
    """The formulated system of the device-resident model (first outer
    iteration), its heads, the existing entry point's ||r||_2 and the active
    mask the model uses (ibound >= 1: CHD cells are excluded)."""
    gs = GwfSolution(case)
    try:
        gs.kstp = 1
        gs.prepare_timestep(1.0, True)
        amat, rhs = gs.formulate(1, want_system=True)
        x = gs.heads()
        l2, _ = gs.residual_norms()
    finally:
        gs.destroy()
    active = np.ones(case.nodes, dtype=np.int32)
    for p in case.packages:
        if p.kind == "CHD":
            active[np.asarray(p.nodes)] = -1
    return amat, rhs, x, active, l2


def _noalloc(case, amat, rhs, x, active):
    fn = _lib.load().mf6hip_csr_residual_norms_noalloc
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    H = HipBuffers()
    try:
        d = [H.put(v) for v in (ia, ja, amat, rhs, x, active)]
        out = H.put(np.full(4, -7.0))
        rc = fn(case.nodes, ja.size, *d, out, None)
        assert rc == 0, _lib.last_error()
        return H.get(out, 4)     # hipMemcpy waits for the null-stream work
    finally:
        H.free()


def _check(case):
    amat, rhs, x, active, l2_ref = _device_system(case)
    got = _noalloc(case, amat, rhs, x, active)
    ia, ja = dis_structure(case.nlay, case.nrow, case.ncol)
    ax = np.zeros(case.nodes)
    O.lib().orc_amux(case.nodes, x, ax, amat, ja - 1, ia - 1)
    r = np.where(active >= 1, ax + (-1.0) * rhs, 0.0)
    k = int(np.argmax(np.abs(r)))
    print(case.name, case.nodes, "l2 noalloc", got[0], "existing", l2_ref, "max r", got[2], r[k])
    assert abs(got[0] - l2_ref) <= 1e-12 * l2_ref
    assert abs(got[1] - np.sqrt(np.sum(np.where(active >= 1, rhs, 0.0) ** 2))) <= \
        1e-12 * got[1]
    assert got[2] == r[k] and int(got[3]) == k + 1


def test_noalloc_residual_norms_match_existing_entry_point_twice():
    _check(configs.steady_small(nlay=3, nrow=60, ncol=64, nwel=10))
    _check(configs.steady_small(nlay=2, nrow=17, ncol=20, nwel=3))


def test_noalloc_rejects_sizes_beyond_its_maximum():
    fn = _lib.load().mf6hip_csr_residual_norms_noalloc
    p = C.c_void_p(1)
    assert fn((1 << 30) + 1, (1 << 30) + 1, p, p, p, p, p, None, p, None) == 1
    assert "2^30" in _lib.last_error()
