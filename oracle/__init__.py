"""CPU oracle of the BMFR frame chain: test infrastructure only (see bmfr_oracle.h)."""
