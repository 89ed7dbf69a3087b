"""ORACLE — test infrastructure only: a CPU restatement (C, with a ctypes wrapper) of the
reference's PISO kernels, Krylov solvers and AMG, used by the tests, the smoke check and the
benchmark's CPU baseline as the checker. The product package never imports it."""
