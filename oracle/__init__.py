"""CPU oracle for the RFP hot path (test infrastructure only; see rfp_oracle)."""
