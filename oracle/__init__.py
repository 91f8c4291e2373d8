"""ORACLE — test infrastructure only (CPU restatement of deepmr's NUFFT).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product path never does.
"""
