"""CPU oracles for the steady-solve hot path -- test infrastructure only.

liboracle.so (joe_oracle.cpp) restates the reference's algorithm; _ref/ holds
the reference's own build (oracle/Makefile, target ref) with a dump harness.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.
"""
