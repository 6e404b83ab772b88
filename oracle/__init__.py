"""CPU oracle package — test infrastructure only (see cpl_oracle.hpp)."""
