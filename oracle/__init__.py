"""CPU oracle for the volumetric transient path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker; the product path
(totrilib_amd) never does.
"""
