"""CPU oracle of the SPD square-root path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the timed CPU baseline); the
product library never does.
"""
