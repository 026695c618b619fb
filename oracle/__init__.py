"""ORACLE -- test infrastructure only (the checker, never the product).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package.

  khp       : ctypes binding of liboracle.so, the C restatement of the
              reference resolver and coherency controller (khp_oracle.c)
  refkhp    : ctypes binding of _ref/libkhpref.so, the reference's own KHP
              engine built from its sources (ref_harness.cc + Makefile)
  pointwise : brute-force per-point semantics on tiny spaces (numpy)
"""
