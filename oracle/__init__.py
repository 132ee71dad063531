"""CPU oracle of the transport hot path -- test infrastructure only."""
