"""CPU FP64 oracle for the HEOM hot path -- test infrastructure only (the checker, never the product)."""
