"""CPU oracle of the TRiSK hot path -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
as the checker; the product (wavetrisk_hydrostatic_amd) never uses it.
"""
