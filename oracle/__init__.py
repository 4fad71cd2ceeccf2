"""CPU oracle of the FUSS explicit step (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package: it is the checker of the HIP path, never part of the product.
"""
