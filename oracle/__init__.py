"""CPU oracle of the OpenSayal per-step solver: test infrastructure only.

Imported by tests/, by ``__graft_entry__.smoke()`` (as the checker) and by
bench.py's CPU-baseline leg; never by the product package.
"""
