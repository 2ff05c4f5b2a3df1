"""CPU oracle of the MPS time step (test infrastructure only: tests/,
__graft_entry__.smoke() and bench.py's CPU baseline load it; the product
library never does)."""
