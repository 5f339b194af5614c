"""CPU oracle for the MODFLOW 6 IMS / GWF hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as
the thing measured or shipped.  The product (``modflow6_amd``) never imports
it.

* ``mf6_oracle.c`` restates the reference routines (file:line cited per
  function) in C; ``oracle.py`` wraps it with ctypes and drives a full DIS
  GWF simulation (formulate, sln_ls row fixes, IMS solve, outer loop).
* ``_ref/libmf6ref.so`` is the reference's own IMS core compiled from
  /root/reference by ``Makefile`` (target ``ref``) and is used only to pin
  the restatement.
"""
