"""CPU oracle of the snow MPM step (test infrastructure only; see mpm_oracle.hpp)."""
