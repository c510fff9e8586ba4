"""Oracle package: the FP64 NumPy restatement of ARL's hot path (test infrastructure only)."""
