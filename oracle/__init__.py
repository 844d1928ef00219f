"""CPU oracle of the NEMO hot path -- test infrastructure only (see nemo_oracle.c)."""
