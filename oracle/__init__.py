"""ORACLE -- test infrastructure only.

FP64 CPU restatement of the reference's RKS B3LYP Fock build and SCF loop
(pyscf_ipu/nanoDFT/nanoDFT.py, sparse_symmetric_ERI.py, exchange_correlation/*,
electron_repulsion/direct.py).  Only `tests/`, `__graft_entry__.smoke()` and
the `cpu_baseline` leg of `bench.py` import this package, and only as the
checker (or the timed CPU baseline).  The product package `pyscf_ipu_amd`
never imports it.

Modules:
  cint        ctypes front of cint_oracle.c (Boys, Rys roots, ERIs, int1e, J/K)
  md_eri      second, independent ERI algorithm (THO / McMurchie-Davidson)
  indices     tril / s8 index arithmetic and the 32-entry J/K index table
  xc          B3LYP (XCFun-derived formulas) with torch-autograd derivatives
  ao          AO values and gradients on grid points
  rks         nanoDFT-semantics SCF driver (DIIS(9), L^-1 H L^-T eigh, ...)

Parity status per function is listed in each module header; anything not
pinned by a closed form, a reference known-answer value or an independent
algorithm says "parity unpinned".
"""
