"""Test infrastructure: the CPU FP64 oracle of the K/R assembly (see oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
the checker.  The product package never imports it.
"""
