"""Small GWF cases of the reference's autotests, shared by the oracle and
device tests: 1-D flow between two CHD cells, test_gwf_newton01.py and the
bump model of test_gwf_newton_under_relaxation.py (golden heads
autotest/data/ex-gwf-bump/results.hds.cmp, copied under tests/golden)."""
import os
import struct

import numpy as np

from modflow6_amd.configs import GwfCase, ImsOptions, Package

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_ulasav(path):
    """One ulasav head record (src/Utilities/InputOutput.f90:924-940)."""
    b = open(path, "rb").read()
    kstp, kper = struct.unpack("<ii", b[:8])
    ncol, nrow, ilay = struct.unpack("<iii", b[40:52])
    return np.frombuffer(b[52:52 + 8 * ncol * nrow], dtype="<f8").reshape(nrow, ncol)


def line_case(ncol=20, k=3.0):
    nodes = ncol
    return GwfCase(
        name="line", nlay=1, nrow=1, ncol=ncol, delr=np.full(ncol, 10.0),
        delc=np.full(1, 10.0), top2d=np.zeros((1, ncol)),
        botm=np.full((1, 1, ncol), -10.0), k11=np.full(nodes, k),
        k33=np.full(nodes, k), icelltype=np.zeros(nodes, np.int32),
        strt=np.zeros(nodes),
        packages=[Package("CHD", np.array([0, ncol - 1]), np.array([1.0, 10.0]))],
        ims=ImsOptions(inner_dvclose=1e-12, inner_rclose=1e-12, outer_dvclose=1e-12),
        periods=[(1.0, 1, True)])


def newton01_case():
    nlay, nrow, ncol = 2, 3, 3
    nodes = nlay * nrow * ncol
    chd_nodes = np.arange(nrow * ncol, 2 * nrow * ncol)
    return GwfCase(
        name="newton01", nlay=nlay, nrow=nrow, ncol=ncol, delr=np.ones(ncol),
        delc=np.ones(nrow), top2d=np.full((nrow, ncol), 20.0),
        botm=np.stack([np.full((nrow, ncol), 10.0), np.zeros((nrow, ncol))]),
        k11=np.full(nodes, 10.0), k33=np.full(nodes, 10.0),
        icelltype=np.ones(nodes, np.int32), strt=np.full(nodes, 7.0),
        packages=[Package("RCH", np.arange(nrow * ncol), np.ones(nrow * ncol)),
                  Package("CHD", chd_nodes, np.full(chd_nodes.size, 7.0))],
        ims=ImsOptions(inner_dvclose=1e-9, outer_dvclose=1e-9, inner_rclose=1e-9,
                       linear_acceleration="BICGSTAB", inner_maximum=100,
                       outer_maximum=100),
        periods=[(1.0, 1, True)], newton=True)


def bump_case():
    nrow = ncol = 51
    botm = np.loadtxt(os.path.join(GOLD, "ex-gwf-bump", "bottom.txt")).reshape(1, nrow, ncol)
    nodes = nrow * ncol
    left = np.arange(nrow) * ncol
    right = left + ncol - 1
    chd = Package("CHD", np.concatenate([left, right]),
                  np.concatenate([np.full(nrow, 7.5), np.full(nrow, 2.5)]))
    # IMS of the reference test (SIMPLE complexity + its overrides)
    ims = ImsOptions(outer_maximum=75, outer_dvclose=1e-8, inner_maximum=100,
                     inner_dvclose=1e-9, inner_rclose=1e-3,
                     linear_acceleration="BICGSTAB", no_ptc="ALL")
    return GwfCase(
        name="bump", nlay=1, nrow=nrow, ncol=ncol, delr=np.full(ncol, 100.0 / ncol),
        delc=np.full(nrow, 100.0 / nrow), top2d=np.full((nrow, ncol), 25.0),
        botm=botm, k11=np.ones(nodes), k33=np.ones(nodes),
        icelltype=np.ones(nodes, np.int32), strt=np.full(nodes, 7.5), packages=[chd],
        ims=ims, periods=[(1.0, 1, True)], newton=True, newton_ur=True)


# NONLINEAR / OPTIONS variants of the bump model; every one converges to the
# same heads (results.hds.cmp), through different outer-iteration paths
BUMP_NONLINEAR = {
    "ptc": dict(no_ptc=None),
    "simple": dict(under_relaxation="SIMPLE", under_relaxation_gamma=0.7),
    "cooley": dict(under_relaxation="COOLEY", under_relaxation_gamma=0.3),
    "dbd": dict(under_relaxation="DBD", under_relaxation_theta=0.7,
                under_relaxation_kappa=0.1, under_relaxation_gamma=0.2,
                under_relaxation_momentum=0.001),
    # the COMPLEX complexity preset (sln_setouter, NumericalSolution.f90
    # :2656-2669) with PTC allowed
    "complex": dict(under_relaxation="DBD", under_relaxation_theta=0.8,
                    under_relaxation_kappa=1e-4, under_relaxation_gamma=0.0,
                    backtracking_number=20, backtracking_tolerance=1.05,
                    backtracking_reduction_factor=0.1,
                    backtracking_residual_limit=0.002, no_ptc=None),
}


def bump_variant(name):
    import dataclasses
    c = bump_case()
    c.ims = dataclasses.replace(c.ims, **BUMP_NONLINEAR[name])
    return c
