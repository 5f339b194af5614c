"""Build helpers: the HIP library (gfx950) and the Fortran ISO_C_BINDING
module.  ``build_all`` always recompiles."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(force=True):
    """Compile modflow6_amd/libmf6hip.so with hipcc --offload-arch=gfx950."""
    csrc = os.path.join(HERE, "csrc")
    if force:
        _run(["make", "-C", csrc, "clean"])
    _run(["make", "-C", csrc, "-j4"])
    return os.path.join(HERE, "libmf6hip.so")


def build_fortran():
    """Compile the ISO_C_BINDING interface module and its example driver with
    flang when it is available (the reference's host language)."""
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None:
        return None
    fdir = os.path.join(HERE, "fortran")
    _run(["make", "-C", fdir, f"FC={fc}"])
    return os.path.join(fdir, "mf6hip_driver")


def build_oracle():
    """Compile the CPU oracle (oracle/liboracle.so) and, where the reference
    sources are present, oracle/_ref/libmf6ref.so (IMS core),
    oracle/_ref/libmf6refgwf.so (GWF packages and the solution's outer-loop
    procedures) and oracle/_ref/libmf6hipplugin.so (the reference-side
    LinearSolverBaseType plugin with its test driver)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "all"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", odir, "ref", "refgwf"])
        # the reference-side LinearSolverBaseType plugin (links libmf6hip.so)
        if os.path.exists(os.path.join(HERE, "libmf6hip.so")):
            _run(["make", "-C", odir, "plugin"])


def build_all():
    build_hip(force=True)
    build_oracle()
    build_fortran()
