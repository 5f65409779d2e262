"""CPU oracle for the acoustic FD / exact adjoint / RTM imaging path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this
directory: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or as the
timed host-CPU baseline, never as the thing measured or shipped.

``oracle.fd`` restates ``openfwi/physics/gpu.py`` op for op in eager PyTorch
on the CPU (same slicing Laplacian, same evaluation order, checkpoint and
recompute loop), generalized to non-square grids, any shot count and any
number of steps.  At the default geometry it reduces to the reference.
"""
