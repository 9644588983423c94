"""Framework interfaces (quafu/algorithms/interface)."""
