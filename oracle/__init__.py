"""Test infrastructure: the CPU oracle the HIP path is checked against.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package.  The product path never calls it.
"""
