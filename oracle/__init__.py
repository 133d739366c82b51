"""CPU oracle for the streaming dBG hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this package, and only as the checker; the product never does.
"""
