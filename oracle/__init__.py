"""CPU oracle of the Flow Lenia step: test infrastructure only (see oracle/flowlenia_oracle.h)."""
