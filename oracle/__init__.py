"""CPU oracle of the WInDS induced-velocity path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this package, as the checker; the engine
in winds-master_amd/ never imports, links or runs anything from it.
"""
