"""FP64 CPU oracle for PHASER's correlation core (test infrastructure only)."""
