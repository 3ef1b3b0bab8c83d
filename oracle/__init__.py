"""CPU oracle of the VELO path -- test infrastructure only (see velo_oracle.h)."""
