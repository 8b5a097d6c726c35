"""CPU FP64 oracle of the reference hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the CPU baseline.  The shipped
library (generic_mobility_interface_class_amd) never imports, links or calls
it.
"""
