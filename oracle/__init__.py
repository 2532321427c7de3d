"""FP64 CPU restatement of the TensorMol-0.1 BP+EE energy/force path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this
module: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it, and only as the checker (or the timed CPU baseline), never as a
fallback for the GPU path.
"""
