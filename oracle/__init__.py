"""CPU oracle for the RGF energy loop (test infrastructure only; see negf_oracle)."""
