"""CPU oracle of the fv3d residual/time-stepping path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it;
the product (fv3d_amd/) never imports it.
"""
