"""CPU oracle of NEUWON's per-timestep advance -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only
as the checker or the CPU baseline.  The product package neuwon_amd never imports it.
"""
