"""CPU oracle: a restatement of the reference algorithm used as the checker
by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.  Never
imported by the tisphi_amd package."""
