"""CPU oracle (checker) for the sweep/exchange path.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this package.
"""
