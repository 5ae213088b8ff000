"""CPU oracle for the basis -> hamiltonian -> Lanczos path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  It wraps oracle/_build/liboracle.so (the C++ restatement of the
reference) and, where built, oracle/_ref/ (the reference's own basis code
compiled from its sources).
"""
from .oracle import *  # noqa: F401,F403
