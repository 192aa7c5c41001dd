"""Test oracle for the CRMod/CRTomo forward + sensitivity path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

* :mod:`oracle.cpu`  -- ctypes binding of ``liboracle.so`` (crt_oracle.c, the C
  restatement of the reference routines).
* :mod:`oracle.ref`  -- ctypes binding of ``_ref/libcrtref.so``: the reference's own
  Fortran routines compiled from its source plus ``ref_driver.f90``; present
  only where the recipe in ``oracle/Makefile`` could run.
"""
