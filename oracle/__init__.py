"""CPU oracle of the N-body stages -- test infrastructure only (see nbody_oracle.h)."""
