"""CPU oracle of the time-stepping path (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker -- never as the thing measured or shipped.
"""
