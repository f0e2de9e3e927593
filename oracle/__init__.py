"""CPU FP64 oracle of the PBF step — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product (pbf-cuda_amd) never does.
"""
