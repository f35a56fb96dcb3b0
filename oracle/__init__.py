"""CPU oracle for the slime-mold step - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only to check results (or as the timed CPU baseline).  The product
library never loads it.  The C restatement lives in slm_oracle.c; this module
is its ctypes binding (liboracle.so, built by oracle/Makefile).
"""
from .pyoracle import *  # noqa: F401,F403
