"""FP64 CPU oracle of SMILI's imaging cost -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, and only as the checker / CPU baseline.
"""
