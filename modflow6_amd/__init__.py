"""modflow6_amd -- MI355X (gfx950) implementation of the MODFLOW 6 IMS hot path.

The product is ``libmf6hip.so`` (HIP kernels behind the C ABI in
``include/mf6hip.h``): the GWF matrix fill (NPF conductance and Newton terms,
STO, WEL / DRN / RCH / CHD) and the IMS linear solve (CG / BiCGSTAB with
ILU0 / MILU0) under the NumericalSolution outer loop.  The Python modules are
the host-side mirror of the reference interfaces used to drive it:

* ``ims``       -- LinearSolverBaseType mirror (HipLinearSolver)
* ``solution``  -- NumericalSolution outer loop over the device system
* ``configs`` / ``fields`` -- synthetic GWF workloads (seeded inputs)
* ``build``     -- compiles the HIP library for gfx950
"""
