"""Oracle: CPU restatement of Covasim's daily transmission step.

TEST INFRASTRUCTURE ONLY -- the checker, never the product. Only tests/,
bench.py's ``cpu_baseline`` leg and ``__graft_entry__.smoke()`` import it.
"""
