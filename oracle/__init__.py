"""CPU oracle for the D2Q9 hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import, call, link or run anything in this directory, and only as the checker
(or the timed CPU baseline); the product path never does.
"""
