"""CPU oracle for the SWARM data back-end GPU stage.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker.
"""
