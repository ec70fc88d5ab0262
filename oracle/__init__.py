"""Test infrastructure: the CPU oracle for the lesgo time step.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product path (lesgo_amd) never does.
"""
