"""Machine-learning data generators (ppafm.ml) on the batched GPU path."""
