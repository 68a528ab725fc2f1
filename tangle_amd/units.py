"""Length units in metres (crates/tangle_python/python/tangle/units.py)."""

m = 1.0
mm = 1.0e-3
um = 1.0e-6
nm = 1.0e-9
