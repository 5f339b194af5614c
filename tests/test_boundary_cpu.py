"""The C-ABI boundary without a GPU: the library loads, exports every entry
point include/mf6hip.h declares, its structs have the layout the header
gives them (ctypes and the Fortran bind(C) mirror against gcc), and argument
validation rejects what the device path does not support before any device
call.  The CSR structure handed over is the reference's DIS structure."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from modflow6_amd import _lib, configs
from modflow6_amd.ims import ImsLinearSettings, dis_structure
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mf6hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(mf6hip_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = header_functions()
    assert len(declared) >= 17
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


STRUCTS = {
    "mf6hip_ims_settings": _lib.ImsSettingsC,
    "mf6hip_cnvg_out": _lib.CnvgOutC,
    "mf6hip_gwf_model": _lib.GwfModelC,
    "mf6hip_ptc": _lib.PtcC,
    "mf6hip_underrelax": _lib.UnderRelaxC,
    "mf6hip_package_options": _lib.PackageOptionsC,
    "mf6hip_nonlinear": _lib.NonlinearC,
    "mf6hip_xmi_iteration": _lib.XmiIterationC,
}
FORTRAN_TYPES = ["mf6hip_ims_settings", "mf6hip_gwf_model", "mf6hip_ptc",
                 "mf6hip_underrelax", "mf6hip_package_options", "mf6hip_nonlinear",
                 "mf6hip_xmi_iteration"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layout_matches_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mf6hip.h"',
             'int main(void) {']
    for sname, cls in STRUCTS.items():
        lines.append(f'printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname} {fname} %zu\\n", offsetof({sname}, {fname}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for sname, cls in STRUCTS.items():
        assert got[(sname, "size")] == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert got[(sname, fname)] == getattr(cls, fname).offset, (sname, fname)


@pytest.mark.skipif(shutil.which("flang") is None, reason="needs flang")
@pytest.mark.parametrize("tname", FORTRAN_TYPES)
def test_fortran_bind_c_types_match_header(tname, tmp_path):
    """The bind(C) derived types of the Fortran module have the C structs'
    size and field offsets (checked through c_loc differences)."""
    cls = STRUCTS[tname]
    mod = os.path.join(ROOT, "modflow6_amd", "fortran", "mf6hip_module.f90")
    prog = tmp_path / "chk.f90"
    fields = [f for f, _ in cls._fields_]
    body = ["program chk", "  use iso_c_binding", "  use mf6hip_module",
            "  implicit none", f"  type({tname}), target :: s",
            f"  type({tname}), target :: a(2)",
            "  integer(c_intptr_t) :: b",
            "  b = transfer(c_loc(s), b)",
            "  print *, 'size', transfer(c_loc(a(2)), b) - transfer(c_loc(a(1)), b)"]
    for f in fields:
        body.append(f"  print *, '{f}', transfer(c_loc(s%{f}), b) - b")
    body.append("end program chk")
    prog.write_text("\n".join(body) + "\n")
    exe = tmp_path / "chk"
    subprocess.check_call(["flang", "-J", str(tmp_path), mod, str(prog), "-o", str(exe)],
                          cwd=tmp_path)
    out = subprocess.check_output([str(exe)], text=True)
    got = {ln.split()[0]: int(ln.split()[1]) for ln in out.strip().split("\n")}
    assert got["size"] == C.sizeof(cls)
    for f in fields:
        assert got[f] == getattr(cls, f).offset, f


def _settings(**kw):
    s = ImsLinearSettings(**kw)
    return s


def _create(settings, nlay=2, nrow=3, ncol=4, ia=None, ja=None, neq=None):
    L = _lib.load()
    ia1, ja1 = dis_structure(nlay, nrow, ncol)
    ia1 = ia1 if ia is None else ia
    ja1 = ja1 if ja is None else ja
    cs = settings.to_c(nlay, nrow, ncol)
    h = C.c_void_p()
    n = nlay * nrow * ncol if neq is None else neq
    rc = L.mf6hip_create(n, ja1.size, _lib.iptr(ia1), _lib.iptr(ja1), C.byref(cs),
                         C.byref(h))
    return rc, h


@pytest.mark.parametrize("kw,code,msg", [
    (dict(iscl=1, preconditioner=2), 3, "SCALING_METHOD"),
    (dict(iord=1), 3, "REORDERING_METHOD"),
    (dict(level=1), 3, "ILUT"),
    (dict(droptol=1e-3), 3, "ILUT"),
    (dict(ilinmeth=3), 1, "LINEAR_ACCELERATION"),
    (dict(relax=1.5), 1, "RELAXATION_FACTOR"),
    (dict(rclose=0.0), 1, "INNER_RCLOSE"),
    (dict(north=-1), 1, "ORTHOGONALIZATIONS"),
    (dict(precision=1, ilinmeth=2, preconditioner=2), 3, "two-level"),
    (dict(kernel_variant=7), 1, "kernel variant"),
])
def test_create_rejects_unsupported_settings(kw, code, msg):
    rc, h = _create(_settings(**kw))
    assert rc == code and not h.value
    assert msg in _lib.last_error()


def test_create_rejects_layer_count_and_structure():
    rc, h = _create(_settings(), nlay=40, nrow=2, ncol=2)
    assert rc == 3 and "32 layers" in _lib.last_error()
    ia, ja = dis_structure(2, 3, 4)
    bad = ja.copy()
    bad[[1, 2]] = bad[[2, 1]]          # columns of row 1 out of order
    rc, h = _create(_settings(), ja=bad)
    assert rc == 1 and "JA does not match" in _lib.last_error()
    rc, h = _create(_settings(), neq=23)
    assert rc == 1 and "NEQ differs" in _lib.last_error()


def test_null_arguments_rejected():
    L = _lib.load()
    assert L.mf6hip_create(1, 1, None, None, None, None) == 1
    assert L.mf6hip_solve(None, 1, 1, None, None, None, None, None, None) != 0
    assert L.mf6hip_gwf_create(None, None) == 1
    assert L.mf6hip_gwf_set_x(None, None) == 1
    assert L.mf6hip_gwf_formulate(None, 1, 1, None, None) == 1
    assert L.mf6hip_gwf_backtrack(None, 0.5, 1e-3, None) == 1
    assert L.mf6hip_gwf_residual_norms(None, None, None) == 1
    assert L.mf6hip_destroy(None) == 0
    buf = C.create_string_buffer(4)
    assert L.mf6hip_last_error(buf, 4) == 0 and len(buf.value) <= 3
    assert L.mf6hip_last_error(None, 0) == 1


@pytest.mark.parametrize("shape", [(1, 1, 7), (2, 3, 4), (3, 5, 5), (1, 51, 51)])
def test_dis_structure_matches_oracle_connections(shape):
    nlay, nrow, ncol = shape
    ia, ja = dis_structure(nlay, nrow, ncol)
    d = O.dis_connections(nlay, nrow, ncol, np.ones(ncol), np.ones(nrow),
                          np.zeros((nrow, ncol)),
                          -np.arange(1, nlay + 1)[:, None, None] * np.ones((nlay, nrow, ncol)))
    assert np.array_equal(ia - 1, d.ia) and np.array_equal(ja - 1, d.ja)


def test_configs_are_deterministic():
    a = configs.steady_small(nlay=2, nrow=20, ncol=20, nwel=5)
    b = configs.steady_small(nlay=2, nrow=20, ncol=20, nwel=5)
    assert np.array_equal(a.k11, b.k11)
    for pa, pb in zip(a.packages, b.packages):
        assert np.array_equal(pa.nodes, pb.nodes) and np.array_equal(pa.v1, pb.v1)
    t = configs.transient_newton(nlay=2, nrow=16, ncol=16, nwel=4)
    assert t.newton and t.has_sto and t.ims.ilinmeth == 2


def test_batch_set_system_checks_structure_and_symmetry():
    """The ensemble's set_system checks its CSR on the host before staging:
    a JA that is not the DIS structure and a nonsymmetric matrix are
    rejected (the batched CG reads every coupling from the upper planes, so
    either would otherwise be solved as a different system); a valid system
    passes the checks and then needs the (device) finalize."""
    from tests.helpers import dis_matrix
    L = _lib.load()
    shape = (2, 3, 4)
    h = C.c_void_p()
    assert L.mf6hip_batch_create(2, C.byref(h)) == 0
    try:
        cs = _settings(precision=1).to_c(*shape)
        assert L.mf6hip_batch_add(h, C.byref(cs)) == 0
        ia, ja, a = dis_matrix(shape, seed=3)
        n = ia.size - 1
        rhs, x0 = np.zeros(n), np.zeros(n)

        def set_sys(ia_, ja_, a_):
            return L.mf6hip_batch_set_system(h, 0, _lib.iptr(ia_), _lib.iptr(ja_), a_.size,
                                             _lib.dptr(a_), _lib.dptr(rhs), _lib.dptr(x0))

        bad = ja.copy()
        bad[[1, 2]] = bad[[2, 1]]
        assert set_sys(ia, bad, a) == 1 and "JA does not match" in _lib.last_error()
        bad_ia = ia.copy()
        bad_ia[5] += 1
        assert set_sys(bad_ia, ja, a) == 1 and "does not match" in _lib.last_error()
        for entry in range(a.size):               # every off-diagonal, one at a time
            r = np.searchsorted(ia - 1, entry, side="right") - 1
            if entry == ia[r] - 1:
                continue
            ns = a.copy()
            ns[entry] += 0.5
            assert set_sys(ia, ja, ns) == 3, entry
            assert "not symmetric" in _lib.last_error()
        assert set_sys(ia, ja, a) == 1 and "not finalized" in _lib.last_error()
    finally:
        L.mf6hip_batch_destroy(h)


def test_xmi_entry_points_validate_without_a_device():
    """The XMI-style entry points check their handle and arguments before any
    device work: NULL handles and settings are rejected."""
    L = _lib.load()
    nl = _lib.NonlinearC(1e-6, 25, 0, 1.0, 1.0, 0.0, 0.0, 0, 0.0, 0.0, 0.0, 1, 0, 1.0, 0.0)
    assert L.mf6hip_xmi_set_nonlinear(None, C.byref(nl)) == 1
    assert L.mf6hip_xmi_prepare_solve(None, 1, 1, 1, 1.0) == 1
    hc = C.c_int(0)
    assert L.mf6hip_xmi_solve(None, C.byref(hc)) == 1
    assert L.mf6hip_xmi_finalize_solve(None, None, None) == 1
    r = _lib.XmiIterationC()
    assert L.mf6hip_xmi_last_iteration(None, C.byref(r)) == 1


@pytest.mark.parametrize("pkg,opt,nbound,rc,msg", [
    (_lib.PKG["RCH"], dict(has_auxmult=1), 0, 3, "AUXMULTNAME"),
    (_lib.PKG["DRN"], dict(iflowred=1, flowred=0.2), 0, 1, "WEL option"),
    (_lib.PKG["WEL"], dict(idrndepth=1), 0, 1, "DRN options"),
    (_lib.PKG["RCH"], dict(icubic_scaling=1), 0, 1, "DRN options"),
    (_lib.PKG["DRN"], dict(idrndepth=1), 3, 1, "needs the depths"),
    (_lib.PKG["WEL"], dict(iflowred=1, flowred=float("nan")), 0, 1, "not a number"),
    (_lib.PKG["WEL"], dict(iflowred=1, flowred=0.2), 0, 1, "no model"),
])
def test_set_package_option_checks(pkg, opt, nbound, rc, msg):
    """Package options are checked before the package data are replaced
    (no handle is needed to see it): AUXMULTNAME is not modelled, options
    of another package type and missing depths are errors; valid options go
    on to the plain package path (here: no handle)."""
    L = _lib.load()
    o = _lib.PackageOptionsC(**opt)
    assert L.mf6hip_gwf_set_package_ex(None, pkg, nbound, None, None, None, C.byref(o)) == rc
    assert msg in _lib.last_error()
    # no options: the plain package path (here: no handle)
    assert L.mf6hip_gwf_set_package_ex(None, pkg, 0, None, None, None, None) == 1
