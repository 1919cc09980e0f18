"""Engine internals: C-ABI binding, plain per-gate autograd Functions, the
circuit lowering + fused pass planner, the adjoint backward, the uneven-batch
runner and the serving loop."""
