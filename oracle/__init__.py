"""CPU oracle for the Barnes-Hut step -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  It is the checker, never the thing measured or shipped.
"""
